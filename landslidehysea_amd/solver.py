"""Host-side mirror of the reference's solver interface over the C-ABI.

The reference drives one simulation through ``shallowWater(...)``
(src/GPU/ShallowWater.cu:125-486); this module exposes the same phases over
``liblhysea_hip.so`` (declared in include/lhysea.h): create, upload, initial
dt, step, download, eta max, gauges and the one-shot driver.  The HIP library
is mandatory: importing works without it, but every call raises
``NativeLibraryMissing`` when the built library is absent.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblhysea_hip.so")

FP64 = 0
FP32C = 1
KERNEL_FOURPASS = 0
KERNEL_FUSED = 1
# status codes and health bits of include/lhysea.h
ERR_GPU_MEMORY = 1
ERR_ARGUMENT = 3
ERR_NUMERIC = 7
HEALTH_NONFINITE = 1
HEALTH_NEGATIVE_DEPTH = 2


class NativeLibraryMissing(RuntimeError):
    pass


class LhyseaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lhysea error {code}: {msg}")
        self.code = code


class CParams(C.Structure):
    """Mirror of ``lhysea_params`` (include/lhysea.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("friction_law", C.c_int32), ("precision", C.c_int32),
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


@dataclass
class Params:
    """The reference's parameter set in internal units (src/GPU/Problema.cxx:226-296).

    ``angle_deg`` is the Coulomb friction angle (one value) or the four
    Pouliquen angles; ``borde`` is (sup, inf, izq, der)."""
    nx: int
    ny: int
    dx: float
    dy: float
    borde: tuple = (-1.0, -1.0, -1.0, -1.0)
    cfl: float = 0.5
    r: float = 0.5
    angle_deg: tuple = (10.0,)
    mfc: float = 0.08
    mf0: float = 0.03
    mfs: float = 0.03
    vmax1: float = 0.0
    vmax2: float = 0.0
    g: float = 9.81
    eps_h: float = 5e-3
    L: float = 1.0
    H: float = 1.0
    friction_law: int = 0
    precision: int = FP64
    peso: float = 1.0
    beta: float = 1.0

    def to_c(self) -> CParams:
        p = CParams()
        p.nx, p.ny, p.friction_law, p.precision = self.nx, self.ny, self.friction_law, self.precision
        p.dx, p.dy = self.dx, self.dy
        p.borde_sup, p.borde_inf, p.borde_izq, p.borde_der = self.borde
        p.cfl, p.r = self.cfl, self.r
        a = tuple(self.angle_deg)
        if len(a) == 1:
            a = a * 4
        p.angle1, p.angle2, p.angle3, p.angle4 = [x * math.pi / 180.0 for x in a]
        p.mfc, p.mf0, p.mfs, p.vmax1, p.vmax2 = self.mfc, self.mf0, self.mfs, self.vmax1, self.vmax2
        p.g, p.eps_h, p.L, p.H = self.g, self.eps_h, self.L, self.H
        p.peso, p.beta = self.peso, self.beta
        return p


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None


class COutputSpec(C.Structure):
    """Mirror of ``lhysea_output_spec`` (include/lhysea.h)."""
    _fields_ = [("prefix", C.c_char_p), ("grid_name", C.c_char_p), ("save_interval", C.c_double),
                ("xmin", C.c_double), ("ymin", C.c_double), ("Hmin", C.c_double), ("Q", C.c_double),
                ("T", C.c_double), ("sim_time", C.c_double), ("gauges", _ip), ("ngauges", C.c_int32)]

# every entry point of include/lhysea.h with its ctypes signature
SIGNATURES = {
    "lhysea_device_probe": (C.c_int, []),
    "lhysea_last_error": (C.c_char_p, []),
    "lhysea_create": (C.c_int, [C.POINTER(CParams), C.POINTER(C.c_void_p)]),
    "lhysea_destroy": (C.c_int, [C.c_void_p]),
    "lhysea_upload": (C.c_int, [C.c_void_p, _dp, _dp]),
    "lhysea_download": (C.c_int, [C.c_void_p, _dp]),
    "lhysea_initial_dt": (C.c_int, [C.c_void_p, _dp]),
    "lhysea_set_dt": (C.c_int, [C.c_void_p, C.c_double]),
    "lhysea_step": (C.c_int, [C.c_void_p, C.c_int]),
    "lhysea_run_until": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _ip]),
    "lhysea_sync": (C.c_int, [C.c_void_p]),
    "lhysea_get_time": (C.c_int, [C.c_void_p, _dp, _dp, _ip]),
    "lhysea_health": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), _ip]),
    "lhysea_track_eta_max": (C.c_int, [C.c_void_p, C.c_int]),
    "lhysea_eta_max": (C.c_int, [C.c_void_p, _dp, _dp]),
    "lhysea_set_gauges": (C.c_int, [C.c_void_p, _ip, C.c_int32, C.c_int32]),
    "lhysea_get_gauges": (C.c_int, [C.c_void_p, _dp, _ip]),
    "lhysea_dt_history": (C.c_int, [C.c_void_p, _dp, C.c_int32, _ip]),
    "lhysea_set_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_set_adjoint_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_time_steps": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    "lhysea_set_tile_split": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_tile_classes": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "lhysea_adjoint_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, C.c_int32, _dp]),
    "lhysea_adjoint_backward": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "lhysea_adjoint_backward_angles": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "lhysea_adjoint": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, C.c_int32, _dp, _dp, _dp, _dp, _dp,
                                 _dp]),
    "lhysea_shallow_water": (C.c_int, [C.POINTER(CParams), _dp, _dp, C.c_double, _dp, _dp, _dp, _ip]),
    "lhysea_output_open": (C.c_int, [C.POINTER(COutputSpec), C.POINTER(CParams), _dp, C.POINTER(C.c_void_p)]),
    "lhysea_output_write": (C.c_int, [C.c_void_p, _dp, C.c_double]),
    "lhysea_output_close": (C.c_int, [C.c_void_p, _dp]),
    "lhysea_shallow_water_output": (C.c_int, [C.POINTER(CParams), _dp, _dp, C.c_double, C.c_int32, C.c_int32,
                                              C.POINTER(COutputSpec), _dp, _ip]),
    "lhysea_step_device": (C.c_int, [C.POINTER(CParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, _dp, _dp,
                                     _ip, C.c_void_p]),
    "lhysea_batch_create": (C.c_int, [C.POINTER(CParams), C.c_int32, C.POINTER(C.c_void_p)]),
    "lhysea_batch_destroy": (C.c_int, [C.c_void_p]),
    "lhysea_batch_size": (C.c_int, [C.c_void_p, _ip]),
    "lhysea_batch_upload": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    "lhysea_batch_download": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "lhysea_batch_initial_dt": (C.c_int, [C.c_void_p]),
    "lhysea_batch_set_stop": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32]),
    "lhysea_batch_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_batch_step_eager": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_batch_sync": (C.c_int, [C.c_void_p]),
    "lhysea_batch_get_time": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _ip]),
    "lhysea_batch_health": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32), _ip]),
    "lhysea_batch_track_eta_max": (C.c_int, [C.c_void_p, C.c_int32, C.c_int]),
    "lhysea_batch_eta_max": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    "lhysea_batch_set_gauges": (C.c_int, [C.c_void_p, C.c_int32, _ip, C.c_int32, C.c_int32]),
    "lhysea_batch_get_gauges": (C.c_int, [C.c_void_p, C.c_int32, _dp, _ip]),
    "lhysea_batch_dt_history": (C.c_int, [C.c_void_p, C.c_int32, _dp, C.c_int32, _ip]),
    "lhysea_batch_tile_shapes": (C.c_int, [_ip, _ip]),
    "lhysea_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "lhysea_batch_get_tile": (C.c_int, [C.c_void_p, _ip]),
    "lhysea_batch_time_steps": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "lhysea_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, _dp, _ip]),
}


def load_library(path: str | None = None):
    """Load liblhysea_hip.so (raises NativeLibraryMissing when it is not built)."""
    global _lib
    if path is None:
        path = LIB_PATH
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise NativeLibraryMissing(f"{path} is not built; run __graft_entry__.build() or "
                                   "python landslidehysea_amd/build.py")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise LhyseaError(rc, load_library().lhysea_last_error().decode())


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def device_probe() -> int:
    return load_library().lhysea_device_probe()


class Simulation:
    """One simulation on the GPU (a handle of the C-ABI)."""

    def __init__(self, params: Params, kernel: int = KERNEL_FUSED):
        self.params = params
        self.lib = load_library()
        self._cp = params.to_c()
        h = C.c_void_p()
        _check(self.lib.lhysea_create(C.byref(self._cp), C.byref(h)))
        self.h = h
        self.N = params.nx * params.ny
        if kernel != KERNEL_FUSED:
            _check(self.lib.lhysea_set_kernel(self.h, kernel))

    def close(self):
        if getattr(self, "h", None):
            self.lib.lhysea_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, W: np.ndarray, H: np.ndarray):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(6, -1)
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
        assert W.shape[1] == self.N and H.size == self.N
        _check(self.lib.lhysea_upload(self.h, _dptr(W), _dptr(H)))

    def download(self) -> np.ndarray:
        W = np.zeros((6, self.params.ny, self.params.nx))
        _check(self.lib.lhysea_download(self.h, _dptr(W)))
        return W

    def initial_dt(self) -> float:
        d = C.c_double()
        _check(self.lib.lhysea_initial_dt(self.h, C.byref(d)))
        return d.value

    def set_dt(self, dt: float):
        _check(self.lib.lhysea_set_dt(self.h, dt))

    def step(self, n: int = 1):
        _check(self.lib.lhysea_step(self.h, n))

    def sync(self):
        _check(self.lib.lhysea_sync(self.h))

    def run_until(self, t_end: float, max_steps: int = 2**31 - 1) -> int:
        """Step while t < t_end (at most max_steps); returns the steps run."""
        n = C.c_int32()
        _check(self.lib.lhysea_run_until(self.h, t_end, max_steps, C.byref(n)))
        return n.value

    def time(self):
        t, d, n = C.c_double(), C.c_double(), C.c_int32()
        _check(self.lib.lhysea_get_time(self.h, C.byref(t), C.byref(d), C.byref(n)))
        return t.value, d.value, n.value

    def health(self):
        """(HEALTH_* bits raised since the last upload, first step that raised one or -1)."""
        f, s = C.c_uint32(), C.c_int32()
        _check(self.lib.lhysea_health(self.h, C.byref(f), C.byref(s)))
        return f.value, s.value

    def track_eta_max(self, on: bool = True):
        _check(self.lib.lhysea_track_eta_max(self.h, 1 if on else 0))

    def eta_max(self):
        v = np.zeros((self.params.ny, self.params.nx))
        t = np.zeros_like(v)
        _check(self.lib.lhysea_eta_max(self.h, _dptr(v), _dptr(t)))
        return v, t

    def set_gauges(self, cells, max_steps: int):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        _check(self.lib.lhysea_set_gauges(self.h, c.ctypes.data_as(_ip), len(c), max_steps))
        self._ng = len(c)
        self._gcap = max_steps

    def gauges(self) -> np.ndarray:
        out = np.zeros((self._gcap, self._ng))
        n = C.c_int32()
        _check(self.lib.lhysea_get_gauges(self.h, _dptr(out), C.byref(n)))
        return out[: n.value]

    def dt_history(self, max_steps: int) -> np.ndarray:
        out = np.zeros(max_steps)
        n = C.c_int32()
        _check(self.lib.lhysea_dt_history(self.h, _dptr(out), max_steps, C.byref(n)))
        return out[: n.value]

    def set_tile_split(self, mode: int):
        """FP32C launch split by tile class: -1 automatic, 0 every tile general, 1 always split."""
        _check(self.lib.lhysea_set_tile_split(self.h, mode))

    def tile_classes(self):
        """(class tiles, general tiles of the next step, last launches split)."""
        t, g, sp = C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.lib.lhysea_tile_classes(self.h, C.byref(t), C.byref(g), C.byref(sp)))
        return t.value, g.value, bool(sp.value)

    def time_steps(self, nsteps: int, per_kernel: bool = True):
        tot, ker = C.c_double(), C.c_double()
        _check(self.lib.lhysea_time_steps(self.h, nsteps, C.byref(tot), C.byref(ker) if per_kernel else None))
        return tot.value, (ker.value if per_kernel else None)


    def set_adjoint_kernel(self, mode: int):
        """Form of the FP32C adjoint's reverse step: KERNEL_FUSED (default) or
        KERNEL_FOURPASS (a cross-check; gradients agree to rounding)."""
        _check(self.lib.lhysea_set_adjoint_kernel(self.h, mode))

    def adjoint_forward(self, nsteps: int, nsnap: int, gauges) -> np.ndarray:
        """Forward pass of the adjoint pair from the uploaded state: returns the
        gauge series eta (nsteps, ng) and keeps the checkpoints in the handle
        until adjoint_backward."""
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        eta = np.zeros((nsteps, len(g)))
        _check(self.lib.lhysea_adjoint_forward(self.h, nsteps, nsnap, g.ctypes.data_as(_ip), len(g), _dptr(eta)))
        self._adj = (nsteps, len(g))
        return eta

    def adjoint_backward(self, seeds):
        """Reverse sweep for the seeds dL/deta (nsteps, ng) of any loss of the
        gauge series: dL/dW0 (6, ny, nx), dL/dangle of the four friction
        angles (radians; Coulomb: only angle1 enters), dL/dangle1, and dL/dmu
        of the Coulomb coefficient (None under the Pouliquen law)."""
        if not getattr(self, "_adj", None):
            raise LhyseaError(3, "no adjoint forward pass to reverse")
        nsteps, ng = self._adj
        seeds = np.ascontiguousarray(seeds, dtype=np.float64).reshape(nsteps, ng)
        gW = np.zeros((6, self.params.ny, self.params.nx))
        stats = np.zeros(6)
        if self.params.friction_law == 1:
            gang4 = np.zeros(4)
            _check(self.lib.lhysea_adjoint_backward_angles(self.h, _dptr(seeds), _dptr(gW), _dptr(gang4),
                                                           _dptr(stats)))
            gmu = None
        else:
            gm, gang = C.c_double(), C.c_double()
            _check(self.lib.lhysea_adjoint_backward(self.h, _dptr(seeds), _dptr(gW), C.byref(gm), C.byref(gang),
                                                    _dptr(stats)))
            gmu, gang4 = gm.value, np.array([gang.value, 0.0, 0.0, 0.0])
        self._adj = None
        return {"grad_W0": gW, "grad_mu": gmu, "grad_angle1": float(gang4[0]), "grad_angles": gang4,
                "forward_ms": stats[0], "reverse_ms": stats[1], "advances": int(stats[2]),
                "reverse_steps": int(stats[3]), "snapshots": int(stats[4]), "optimal_advances": int(stats[5])}

    def adjoint(self, nsteps: int, nsnap: int, gauges, obs):
        """Gauge-misfit loss L = sum (eta - obs)^2 over nsteps steps from the
        uploaded state and its gradients: dL/dW0 (6, ny, nx), dL/dmu (Coulomb),
        dL/dangle1, and dL/dangle of the four angles (radians)."""
        if self.params.friction_law == 1:
            # the forward / backward pair with the misfit's seeds 2 (eta - obs)
            eta = self.adjoint_forward(nsteps, nsnap, gauges)
            d = eta - np.asarray(obs, dtype=np.float64).reshape(eta.shape)
            out = self.adjoint_backward(2.0 * d)
            out["loss"] = float(np.sum(d * d))
            return out
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(nsteps, len(g))
        gW = np.zeros((6, self.params.ny, self.params.nx))
        loss, gmu, gang = C.c_double(), C.c_double(), C.c_double()
        stats = np.zeros(6)
        _check(self.lib.lhysea_adjoint(self.h, nsteps, nsnap, g.ctypes.data_as(_ip), len(g), _dptr(obs),
                                       C.byref(loss), _dptr(gW), C.byref(gmu), C.byref(gang), _dptr(stats)))
        return {"loss": loss.value, "grad_W0": gW, "grad_mu": gmu.value, "grad_angle1": gang.value,
                "grad_angles": np.array([gang.value, 0.0, 0.0, 0.0]), "forward_ms": stats[0], "reverse_ms": stats[1], "advances": int(stats[2]),
                "reverse_steps": int(stats[3]), "snapshots": int(stats[4]), "optimal_advances": int(stats[5])}


def batch_tile_shapes() -> list:
    """(TX, TY, threads) of each tile shape the batched kernel offers."""
    lib = load_library()
    n = C.c_int32()
    _check(lib.lhysea_batch_tile_shapes(None, C.byref(n)))
    a = np.zeros(3 * n.value, dtype=np.int32)
    _check(lib.lhysea_batch_tile_shapes(a.ctypes.data_as(_ip), C.byref(n)))
    return [tuple(int(v) for v in a[3 * i:3 * i + 3]) for i in range(n.value)]


class Batch:
    """An ensemble of independent FP32C scenarios stepped together (a batch
    handle of the C-ABI).  Scenario ``sid`` offers the calls of Simulation."""

    def __init__(self, params: list):
        self.params = list(params)
        self.lib = load_library()
        arr = (CParams * len(self.params))(*[p.to_c() for p in self.params])
        self._cp = arr
        h = C.c_void_p()
        _check(self.lib.lhysea_batch_create(arr, len(self.params), C.byref(h)))
        self.h = h
        self._gshape = {}

    def __len__(self):
        return len(self.params)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lhysea_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, sid: int, W: np.ndarray, H: np.ndarray):
        p = self.params[sid]
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(6, -1)
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
        assert W.shape[1] == p.nx * p.ny and H.size == p.nx * p.ny
        _check(self.lib.lhysea_batch_upload(self.h, sid, _dptr(W), _dptr(H)))

    def download(self, sid: int) -> np.ndarray:
        p = self.params[sid]
        W = np.zeros((6, p.ny, p.nx))
        _check(self.lib.lhysea_batch_download(self.h, sid, _dptr(W)))
        return W

    def initial_dt(self):
        _check(self.lib.lhysea_batch_initial_dt(self.h))

    def set_stop(self, sid: int, t_end: float = math.inf, max_steps: int = -1):
        _check(self.lib.lhysea_batch_set_stop(self.h, sid, t_end, max_steps))

    def step(self, n: int = 1, graph: bool = True):
        fn = self.lib.lhysea_batch_step if graph else self.lib.lhysea_batch_step_eager
        _check(fn(self.h, n))

    def sync(self):
        _check(self.lib.lhysea_batch_sync(self.h))

    def time(self, sid: int):
        t, d, n = C.c_double(), C.c_double(), C.c_int32()
        _check(self.lib.lhysea_batch_get_time(self.h, sid, C.byref(t), C.byref(d), C.byref(n)))
        return t.value, d.value, n.value

    def health(self, sid: int):
        """Simulation.health() of scenario sid."""
        f, s = C.c_uint32(), C.c_int32()
        _check(self.lib.lhysea_batch_health(self.h, sid, C.byref(f), C.byref(s)))
        return f.value, s.value

    def track_eta_max(self, sid: int, on: bool = True):
        _check(self.lib.lhysea_batch_track_eta_max(self.h, sid, 1 if on else 0))

    def eta_max(self, sid: int):
        p = self.params[sid]
        v = np.zeros((p.ny, p.nx))
        t = np.zeros_like(v)
        _check(self.lib.lhysea_batch_eta_max(self.h, sid, _dptr(v), _dptr(t)))
        return v, t

    def set_gauges(self, sid: int, cells, max_steps: int):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        _check(self.lib.lhysea_batch_set_gauges(self.h, sid, c.ctypes.data_as(_ip), len(c), max_steps))
        self._gshape[sid] = (max_steps, len(c))

    def gauges(self, sid: int) -> np.ndarray:
        out = np.zeros(self._gshape.get(sid, (0, 0)))
        n = C.c_int32()
        _check(self.lib.lhysea_batch_get_gauges(self.h, sid, _dptr(out) if out.size else None, C.byref(n)))
        return out[: n.value]

    def dt_history(self, sid: int, max_steps: int) -> np.ndarray:
        out = np.zeros(max_steps)
        n = C.c_int32()
        _check(self.lib.lhysea_batch_dt_history(self.h, sid, _dptr(out), max_steps, C.byref(n)))
        return out[: n.value]

    def set_tile(self, tile: int):
        _check(self.lib.lhysea_batch_set_tile(self.h, tile))

    def tile(self) -> int:
        t = C.c_int32()
        _check(self.lib.lhysea_batch_get_tile(self.h, C.byref(t)))
        return t.value

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(self.lib.lhysea_batch_time_steps(self.h, nsteps, C.byref(ms)))
        return ms.value

    def autotune(self, nsteps: int = 20):
        """Time every tile shape and keep the fastest; returns (best, ms per shape)."""
        n = len(batch_tile_shapes())
        ms = np.zeros(n)
        best = C.c_int32()
        _check(self.lib.lhysea_batch_autotune(self.h, nsteps, _dptr(ms), C.byref(best)))
        return best.value, ms


def shallow_water(params: Params, W: np.ndarray, H: np.ndarray, t_end: float):
    """One-shot driver with shallowWater's contract; returns (W, eta_max, eta_max_time, seconds, steps)."""
    lib = load_library()
    cp = params.to_c()
    W = np.ascontiguousarray(W, dtype=np.float64).copy()
    H = np.ascontiguousarray(H, dtype=np.float64)
    em = np.zeros(params.nx * params.ny)
    et = np.zeros_like(em)
    el = C.c_double()
    n = C.c_int32()
    _check(lib.lhysea_shallow_water(C.byref(cp), _dptr(W), _dptr(H), t_end, _dptr(em), _dptr(et),
                                    C.byref(el), C.byref(n)))
    return W, em.reshape(params.ny, params.nx), et.reshape(params.ny, params.nx), el.value, n.value


def output_spec(prefix: str, save_interval: float = -1.0, grid_name: str = "", xmin: float = 0.0,
                ymin: float = 0.0, Hmin: float = 0.0, Q: float = 1.0, T: float = 1.0, sim_time: float = 0.0,
                gauges=None):
    """lhysea_output_spec; returns (spec, keep-alive) — the gauge array must outlive the calls."""
    g = None if gauges is None else np.ascontiguousarray(gauges, dtype=np.int32)
    spec = COutputSpec(prefix.encode(), grid_name.encode(), save_interval, xmin, ymin, Hmin, Q, T, sim_time,
                       None if g is None else g.ctypes.data_as(_ip), 0 if g is None else len(g))
    return spec, g


class Output:
    """Output files of a run (lhysea_output_open/write/close); host code, no GPU."""

    def __init__(self, params: Params, H: np.ndarray, prefix: str, **spec_kw):
        self.lib = load_library()
        self._cp = params.to_c()
        self._spec, self._keep = output_spec(prefix, **spec_kw)
        self._H = np.ascontiguousarray(H, dtype=np.float64)
        h = C.c_void_p()
        _check(self.lib.lhysea_output_open(C.byref(self._spec), C.byref(self._cp), _dptr(self._H), C.byref(h)))
        self.h = h

    def write(self, W: np.ndarray, t: float):
        _check(self.lib.lhysea_output_write(self.h, _dptr(np.ascontiguousarray(W, dtype=np.float64)), t))

    def close(self, eta_max: np.ndarray | None = None):
        if self.h:
            em = None if eta_max is None else np.ascontiguousarray(eta_max, dtype=np.float64)
            h, self.h = self.h, None
            _check(self.lib.lhysea_output_close(h, None if em is None else _dptr(em)))


def shallow_water_output(params: Params, W: np.ndarray, H: np.ndarray, t_end: float, prefix: str,
                         max_steps: int = 2**31 - 1, verbose: bool = False, **spec_kw):
    """lhysea_shallow_water_output: the run with shallowWater's output files;
    returns (final W, seconds, steps)."""
    lib = load_library()
    cp = params.to_c()
    spec, keep = output_spec(prefix, **spec_kw)
    W = np.ascontiguousarray(W, dtype=np.float64).copy()
    H = np.ascontiguousarray(H, dtype=np.float64)
    el, n = C.c_double(), C.c_int32()
    _check(lib.lhysea_shallow_water_output(C.byref(cp), _dptr(W), _dptr(H), t_end, max_steps, 1 if verbose else 0,
                                           C.byref(spec), C.byref(el), C.byref(n)))
    del keep
    return W, el.value, n.value


def step_device(params: Params, H_dev: int, W_dev: int, W_work_dev: int, nsteps: int, t: float = 0.0,
                dt: float = 0.0, stream: int | None = None):
    """Steps of a state in caller-owned device memory without allocation
    (lhysea_step_device): device addresses of the depth, the state and a
    second state buffer of its layout.  Returns (t, dt_next, result_in_work)."""
    lib = load_library()
    cp = params.to_c()
    tt, dd, w = C.c_double(t), C.c_double(dt), C.c_int32()
    _check(lib.lhysea_step_device(C.byref(cp), C.c_void_p(H_dev), C.c_void_p(W_dev), C.c_void_p(W_work_dev),
                                  int(nsteps), C.byref(tt), C.byref(dd), C.byref(w),
                                  C.c_void_p(stream) if stream else None))
    return tt.value, dd.value, w.value
