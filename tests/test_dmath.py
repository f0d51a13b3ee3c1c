"""The FP64 step's deterministic exp / log / cbrt (CPU only).

The FP64 kernels evaluate exp and powf of the reference
(src/GPU/Volumen_kernel.cu:15, :68, :104; src/GPU/Arista_kernel.cu:40, :136)
with functions built from IEEE-exact operations only
(landslidehysea_amd/csrc/lhysea_dmath.hpp), and the oracle with its own C
restatement of them (lh_exp, lh_log, lh_cbrt in oracle/lhysea_oracle.c).
Checks: the header compiled for the host equals the oracle's functions bit
for bit on a million inputs per function over the ranges the step feeds them
(and beyond), and both stay within a few ulp of the C library; special
values behave as documented.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "dmath_exports.cpp")


@pytest.fixture(scope="module")
def dm(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("dmath") / "libdmath.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-w",
                    "-I", os.path.join(ROOT, "tests", "native"), SRC, "-o", so], check=True)
    L = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    L.dm_eval.argtypes = [C.c_int, dp, dp, dp, C.c_long]

    def ev(fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(x if y is None else y, dtype=np.float64)
        out = np.empty_like(x)
        L.dm_eval(fn, x.ctypes.data_as(dp), y.ctypes.data_as(dp), out.ctypes.data_as(dp), x.size)
        return out
    return ev


def oracle_fn(oracle, name):
    f = getattr(oracle.lib(), name)
    f.argtypes = [C.c_double]
    f.restype = C.c_double
    return np.vectorize(f, otypes=[np.float64])


def ulps(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.spacing(np.abs(b))


def inputs(rng):
    return {
        "exp": np.concatenate([rng.uniform(-707.0, 709.0, 300000), rng.uniform(-1.0, 1.0, 300000),
                               -np.exp(rng.uniform(-40.0, 6.5, 400000))]),
        "log": np.exp(rng.uniform(-700.0, 700.0, 1000000)),
        "cbrt": np.concatenate([np.exp(rng.uniform(-700.0, 700.0, 600000)), rng.uniform(0.0, 1e3, 400000)]),
    }


def test_bitwise_equal_to_oracle_and_close_to_libm(dm, oracle):
    rng = np.random.default_rng(2024)
    X = inputs(rng)
    for fn, name, ref, lim in ((0, "lh_exp", np.exp, 2.0), (1, "lh_log", np.log, 2.0), (2, "lh_cbrt", np.cbrt, 1.0)):
        x = X[("exp", "log", "cbrt")[fn]]
        got = dm(fn, x)
        sub = x[::50]                                   # the oracle through ctypes: a 2% sample
        assert np.array_equal(got[::50], oracle_fn(oracle, name)(sub)), name
        err = ulps(got, ref(x))
        assert err.max() <= lim, (name, float(err.max()))


def test_powers(dm):
    rng = np.random.default_rng(7)
    h = np.concatenate([np.exp(rng.uniform(-30.0, 8.0, 200000)), rng.uniform(0.0, 500.0, 200000)])
    # against h cbrt(h): h ** (4.0 / 3.0) raises to fl(4/3), off by |ln h| 7e-17 relative
    assert ulps(dm(3, h), h * np.cbrt(h)).max() <= 3.0
    x = np.exp(rng.uniform(-20.0, 20.0, 200000))
    y = rng.uniform(-3.0, 3.0, 200000)
    assert ulps(dm(4, x, y), x ** y).max() <= 64.0            # exp(y log x): log's error times |y log x|
    assert dm(3, np.array([0.0]))[0] == 0.0


def test_special_values(dm):
    assert dm(0, np.array([0.0]))[0] == 1.0
    assert dm(0, np.array([-800.0, -707.5]))[0] == 0.0 and dm(0, np.array([-707.5]))[0] == 0.0
    assert np.isinf(dm(0, np.array([710.0]))[0])
    assert math.isnan(dm(0, np.array([np.nan]))[0])
    assert dm(0, np.array([-706.9]))[0] >= np.finfo(np.float64).tiny     # nonzero results are normal
    assert dm(1, np.array([1.0]))[0] == 0.0
    assert np.isneginf(dm(1, np.array([0.0]))[0]) and math.isnan(dm(1, np.array([-1.0]))[0])
    assert dm(2, np.array([-27.0]))[0] == -3.0 and dm(2, np.array([8.0]))[0] == 2.0
