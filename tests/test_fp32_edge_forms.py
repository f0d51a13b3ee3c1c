"""CPU check that the FP32 no-sediment forms equal the general edge and cell update.

The fused FP32C kernel sends a wavefront of edges whose cells all carry water
only (h2 = 0, q2 = 0) through edge1(), and a wavefront of such cells with zero
layer-2 flux sums through cell_update1(): forms of edge() and cell_update()
with the terms that such edges and cells make zero removed.
tests/native/edge1_check.cpp compiles the device functions for the host (g++,
no FMA contraction, so both forms round the same operations) and compares
them on random inputs: layer-1 results equal, layer-2 results zero.
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
    assert out.stdout.count("mismatches 0") == 2
    wet = int(out.stdout.split("wet ")[1].split()[0])
    assert wet > 300000
