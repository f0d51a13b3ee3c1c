"""Register and LDS budget of the FP32C sediment-free kernels (CPU: reads the
AMD assembly of this repository's own gfx950 build).

These kernels are bound by VALU issue and need four 512-thread workgroups
per CU (eight waves per SIMD) to keep the VALU busy.  On gfx950 that holds
only while a wave needs at most 64 VGPRs and at most 80 SGPRs (74 plus the
reserved ones; a wave needing more is allocated 96) and a workgroup at most
40 KB of LDS; crossing the SGPR line alone cost 12% per 4096 x 4096 step on
the MI355X.  The test compiles lhysea_fp32c.hip for the device only and
checks every sediment-free kernel against that budget, with no scratch.
"""
import os
import re
import shutil
import subprocess

import pytest

from landslidehysea_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = ("k_fused_bare", "k_fused_bare_a64", "k_batch_bare", "k_batch_bare_a64")


def kernel_resources(asm: str) -> dict:
    """Per-kernel (.amdhsa_next_free_vgpr, _sgpr, group and private segment)."""
    out = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", asm, flags=re.S):
        body = m.group(2)

        def field(name):
            return int(re.search(r"\.amdhsa_" + name + r"\s+(\d+)", body).group(1))

        out[m.group(1)] = dict(vgpr=field("next_free_vgpr"), sgpr=field("next_free_sgpr"),
                               lds=field("group_segment_fixed_size"), scratch=field("private_segment_fixed_size"))
    return out


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs hipcc")
def test_sediment_free_kernels_keep_four_workgroups_per_cu(tmp_path):
    asm = tmp_path / "fp32c.s"
    cmd = [B.hipcc(), "--cuda-device-only", "-S"] + [f for f in B.HIP_FLAGS if f not in ("-fPIC", "-shared")] + \
          [os.path.join(B.CSRC, "lhysea_fp32c.hip"), "-o", str(asm)]
    subprocess.run(cmd, check=True, capture_output=True)
    res = kernel_resources(asm.read_text())
    found = {}
    for name, r in res.items():
        for k in KERNELS:
            # mangled name: ...<len>k_fused_bare I... (the a64 forms have their own names)
            if re.search(r"\d+" + k + r"I", name):
                found.setdefault(k, []).append((name, r))
    assert sorted(found) == sorted(KERNELS), sorted(found)
    for k, items in found.items():
        for name, r in items:
            print(k, r)
            assert r["vgpr"] <= 64, (name, r)
            assert r["sgpr"] <= 74, (name, r)
            assert r["lds"] <= 40 * 1024, (name, r)
            assert r["scratch"] == 0, (name, r)
