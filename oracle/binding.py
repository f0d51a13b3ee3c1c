"""ctypes binding of the FP64 oracle (oracle/liblhysea_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (to
check the library's results) and by bench.py's cpu_baseline leg.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblhysea_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class LhParams(C.Structure):
    """Mirror of lh_params in oracle/lhysea_oracle.c."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("friction_law", C.c_int), ("pad_", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("borde_sup", C.c_double), ("borde_inf", C.c_double),
        ("borde_izq", C.c_double), ("borde_der", C.c_double),
        ("cfl", C.c_double), ("r", C.c_double),
        ("angle1", C.c_double), ("angle2", C.c_double), ("angle3", C.c_double), ("angle4", C.c_double),
        ("mfc", C.c_double), ("mf0", C.c_double), ("mfs", C.c_double),
        ("vmax1", C.c_double), ("vmax2", C.c_double),
        ("g", C.c_double), ("eps_h", C.c_double), ("L", C.c_double), ("H", C.c_double),
        ("peso", C.c_double), ("beta", C.c_double),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "lhysea_oracle.c")):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = open_library(LIB_PATH)
    return _lib


def open_library(path: str):
    """Load an oracle build at path with its entry points typed (tests load
    variants of the oracle with a planted mistake this way)."""
    L = C.CDLL(path)
    P = C.POINTER(LhParams)
    L.lh_edge.argtypes = [P, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double, _dp, _dp, C.c_int]
    L.lh_edge.restype = C.c_int
    L.lh_edge_dt0.argtypes = [P, _dp, _dp, C.c_double, C.c_double, C.c_double]
    L.lh_edge_dt0.restype = C.c_double
    L.lh_eig.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.lh_flux1d.argtypes = [_dp, C.c_double, _dp]
    L.lh_pvm_coeffs.argtypes = [C.c_double] * 7 + [_dp]
    L.lh_friction.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double]
    L.lh_friction.restype = C.c_double
    L.lh_filtro.argtypes = [P, _dp, _dp, C.c_double]
    L.lh_dis_implicita.argtypes = [P, _dp, _dp, _dp, _dp, C.c_double]
    L.lh_coulomb.argtypes = [P, _dp, _dp, C.c_double]
    L.lh_cell_update.argtypes = [P, _dp, _dp, C.c_double, _dp]
    L.lh_cell_update.restype = C.c_double
    L.lh_oracle_initial_dt.argtypes = [P, _dp, _dp]
    L.lh_oracle_initial_dt.restype = C.c_double
    L.lh_oracle_step.argtypes = [P, _dp, _dp, C.c_double, _dp]
    L.lh_oracle_step.restype = C.c_double
    L.lh_oracle_run.argtypes = [P, _dp, _dp, C.c_int, _dp, _dp, _dp, _ip, C.c_int, _dp, _dp]
    L.lh_oracle_run.restype = C.c_int
    L.lh_oracle_run_fixed_dt.argtypes = [P, _dp, _dp, C.c_int, _dp, _ip, C.c_int, _dp]
    L.lh_oracle_run_fixed_dt.restype = C.c_int
    L.lh_oracle_set_threads.argtypes = [C.c_int]
    L.lh_oracle_get_threads.restype = C.c_int
    L.lh_oracle_wtime.restype = C.c_double
    return L


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def make_params(nx, ny, dx, dy, *, borde=(-1.0, -1.0, -1.0, -1.0), cfl=0.5, r=0.5, angle_deg=10.0,
                mfc=0.08, mf0=0.03, mfs=0.03, vmax1=0.0, vmax2=0.0, g=9.81, eps_h=5e-3, L=1.0, H=1.0,
                friction_law=0, angles_deg=None) -> LhParams:
    p = LhParams()
    p.nx, p.ny, p.friction_law = nx, ny, friction_law
    p.dx, p.dy = dx, dy
    p.borde_sup, p.borde_inf, p.borde_izq, p.borde_der = borde
    p.cfl, p.r = cfl, r
    a = angles_deg if angles_deg is not None else (angle_deg, angle_deg, angle_deg, angle_deg)
    p.angle1, p.angle2, p.angle3, p.angle4 = [x * math.pi / 180.0 for x in a]
    p.mfc, p.mf0, p.mfs, p.vmax1, p.vmax2 = mfc, mf0, mfs, vmax1, vmax2
    p.g, p.eps_h, p.L, p.H = g, eps_h, L, H
    p.peso = p.beta = 1.0
    return p


def set_threads(n: int):
    lib().lh_oracle_set_threads(int(n))


def initial_dt(p, W, H) -> float:
    W = np.ascontiguousarray(W, dtype=np.float64)
    H = np.ascontiguousarray(H, dtype=np.float64)
    return lib().lh_oracle_initial_dt(C.byref(p), ptr(W), ptr(H))


def step(p, W, H, dt, want_acc=False):
    """One step in place on W (6, ny, nx); returns (dt_next, acc or None)."""
    acc = np.zeros((7,) + W.shape[1:]) if want_acc else None
    d = lib().lh_oracle_step(C.byref(p), ptr(W), ptr(H), dt, ptr(acc))
    return d, acc


def run(p, W, H, nsteps, dt=0.0, t=0.0, eta_max=None, gauges=None, record=True):
    """Run nsteps in place on W.  Returns dict with dt_next, t, dt_hist, gauge series."""
    g = np.ascontiguousarray(gauges if gauges is not None else np.zeros(0), dtype=np.int32)
    ng = len(g)
    series = np.zeros((nsteps, ng)) if (record and ng) else None
    hist = np.zeros(nsteps) if record else None
    dti = C.c_double(dt)
    ti = C.c_double(t)
    rc = lib().lh_oracle_run(C.byref(p), ptr(W), ptr(H), nsteps, C.byref(dti), C.byref(ti),
                             ptr(eta_max), ptr(g) if ng else None, ng, ptr(series), ptr(hist))
    if rc != 0:
        raise RuntimeError(f"oracle run failed rc={rc}")
    return {"dt_next": dti.value, "t": ti.value, "dt_hist": hist, "gauges": series}


def run_fixed_dt(p, W, H, dt_seq, gauges):
    g = np.ascontiguousarray(gauges, dtype=np.int32)
    dt_seq = np.ascontiguousarray(dt_seq, dtype=np.float64)
    series = np.zeros((len(dt_seq), len(g)))
    rc = lib().lh_oracle_run_fixed_dt(C.byref(p), ptr(W), ptr(H), len(dt_seq), ptr(dt_seq),
                                      ptr(g), len(g), ptr(series))
    if rc != 0:
        raise RuntimeError("oracle run failed")
    return series


def eig(f, r, g, eps_h, flag):
    f = np.ascontiguousarray(f, dtype=np.float64)
    D = np.zeros(4)
    lib().lh_eig(ptr(f), r, g, eps_h, flag, ptr(D))
    return D


def flux1d(w, eps_h):
    w = np.ascontiguousarray(w, dtype=np.float64)
    F = np.zeros(4)
    lib().lh_flux1d(ptr(w), eps_h, ptr(F))
    return F


def pvm_coeffs(aut1, aut2, aut3, h1ij, h2ij, maxl, eps_h):
    a = np.zeros(3)
    lib().lh_pvm_coeffs(aut1, aut2, aut3, h1ij, h2ij, maxl, eps_h, ptr(a))
    return a


def edge(p, W0, W1, H0, H1, nxv, nyv, length, dt, acc0=None, acc1=None, interior=True):
    W0 = np.ascontiguousarray(W0, dtype=np.float64)
    W1 = np.ascontiguousarray(W1, dtype=np.float64)
    a0 = np.zeros(7) if acc0 is None else np.array(acc0, dtype=np.float64)
    a1 = np.zeros(7) if acc1 is None else np.array(acc1, dtype=np.float64)
    wet = lib().lh_edge(C.byref(p), ptr(W0), ptr(W1), H0, H1, nxv, nyv, length, dt, ptr(a0), ptr(a1),
                        1 if interior else 0)
    return wet, a0, a1


def cell_update(p, Wold, acc, dt):
    Wold = np.ascontiguousarray(Wold, dtype=np.float64)
    acc = np.ascontiguousarray(acc, dtype=np.float64)
    Wn = np.zeros(6)
    d = lib().lh_cell_update(C.byref(p), ptr(Wold), ptr(acc), dt, ptr(Wn))
    return Wn, d


def coulomb(p, n1, n2, dt):
    n1 = np.array(n1, dtype=np.float64)
    n2 = np.array(n2, dtype=np.float64)
    lib().lh_coulomb(C.byref(p), ptr(n1), ptr(n2), dt)
    return n1, n2


def dis_implicita(p, o1, o2, n1, n2, dt):
    o1 = np.array(o1, dtype=np.float64)
    o2 = np.array(o2, dtype=np.float64)
    n1 = np.array(n1, dtype=np.float64)
    n2 = np.array(n2, dtype=np.float64)
    lib().lh_dis_implicita(C.byref(p), ptr(o1), ptr(o2), ptr(n1), ptr(n2), dt)
    return n1, n2


def filtro(p, n1, n2, dt):
    n1 = np.array(n1, dtype=np.float64)
    n2 = np.array(n2, dtype=np.float64)
    lib().lh_filtro(C.byref(p), ptr(n1), ptr(n2), dt)
    return n1, n2


def wtime() -> float:
    return lib().lh_oracle_wtime()
