"""CPU checks that the no-sediment forms equal the general edge, cell update and adjoints.

The fused FP32C kernel sends a wavefront of edges whose cells all carry water
only (h2 = 0, q2 = 0) through edge1(), and a wavefront of such cells with zero
layer-2 flux sums through cell_update1(): forms of edge() and cell_update()
with the terms that such edges and cells make zero removed.
tests/native/edge1_check.cpp compiles the device functions for the host (g++,
no FMA contraction, so both forms round the same operations) and compares
them on random inputs: layer-1 results equal, layer-2 results zero.  The
edges take axis-aligned normals (x-edge (sn, 0), y-edge (0, sn)); an x-edge
and the y-edge of the same states rotated by 90 degrees must give the
rotated fluxes with equal values, for edge() (with sediment too) and edge1()
and both normal signs (a sign or component slip in one axis's projection or
rotation fails it).  The FP64 fused kernel's edge1() and cell_update1() are
checked the same way against the FP64 edge() and cell_update(), and the FP64
edges pass the same rotation check.
tests/native/adjoint_ns_check.cpp does the same for the generated adjoint
functions adj_edge_ns / adj_cell_ns (values and the adjoints of every input,
FP32 and FP64).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_edge1_equals_edge_without_sediment(tmp_path):
    exe = tmp_path / "edge1_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(HERE, "native"),
                    os.path.join(HERE, "native", "edge1_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "400000"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "edges" in out.stdout and "cells" in out.stdout
    assert out.stdout.count("mismatches 0") == 6
    wet = int(out.stdout.split("wet ")[1].split()[0])
    assert wet > 300000
    rotated = int(out.stdout.split("rotated edges")[1].split("wet ")[1].split()[0])
    assert rotated > 300000


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_adjoint_specialisations_equal_general(tmp_path):
    exe = tmp_path / "adjoint_ns_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(HERE, "native"),
                    os.path.join(HERE, "native", "adjoint_ns_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "100000"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert out.stdout.count("mismatches 0") == 4
