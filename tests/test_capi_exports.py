"""The C-ABI library loads and exports every entry point declared in include/lhysea.h (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lhysea.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lhysea_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("lhysea_device_probe", "lhysea_shallow_water", "lhysea_create", "lhysea_step", "lhysea_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from landslidehysea_amd import solver
    lib = solver.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared entry points
    assert sorted(solver.SIGNATURES) == declared()


def test_missing_library_fails_loudly(tmp_path):
    from landslidehysea_amd import solver
    with pytest.raises(solver.NativeLibraryMissing):
        solver.load_library(str(tmp_path / "absent.so"))


def test_probe_without_gpu_reports_no_device():
    from landslidehysea_amd import solver
    rc = solver.device_probe()
    assert rc in (0, 1, 2)


def test_batch_rejects_invalid_scenarios_before_touching_a_device():
    """Argument checks of lhysea_batch_create run before any device call."""
    from landslidehysea_amd import solver
    p = solver.Params(nx=8, ny=8, dx=1.0, dy=1.0, precision=solver.FP64)
    with pytest.raises(solver.LhyseaError) as e:
        solver.Batch([p])
    assert e.value.code == 3 and "FP32C" in str(e.value)
    q = solver.Params(nx=8, ny=8, dx=1.0, dy=1.0, precision=solver.FP32C, friction_law=2)
    with pytest.raises(solver.LhyseaError) as e:
        solver.Batch([q])
    assert e.value.code == 3
    assert len(solver.batch_tile_shapes()) >= 2


@pytest.mark.parametrize("precision", [0, 1])
def test_create_rejects_unknown_friction_law_before_touching_a_device(precision):
    """friction_law outside {0, 1} is an argument error in both precisions
    (no silent Pouliquen for other nonzero values)."""
    from landslidehysea_amd import solver
    for law in (2, -1):
        p = solver.Params(nx=8, ny=8, dx=1.0, dy=1.0, precision=precision, friction_law=law)
        with pytest.raises(solver.LhyseaError) as e:
            solver.Simulation(p)
        assert e.value.code == 3 and "friction_law" in str(e.value)
