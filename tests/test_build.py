"""Build-level guards for the gfx950 kernels (CPU only: hipcc cross-compiles without a GPU).

The fine-level kernels pipeline their neighbour loads; small source changes can make the compiler
spill that pipeline to scratch memory, or place a small array indexed with a run-time value there,
which slows a kernel several times over without changing its results.  This test compiles
kernels.hip for gfx950 and requires zero VGPR spills and zero scratch in every kernel.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = os.path.join(ROOT, "homo3d_amd", "csrc", "kernels.hip")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_register_spills(tmp_path):
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-c", "--offload-device-only",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Rpass-analysis=kernel-resource-usage", KERNELS, "-o", str(tmp_path / "k.o")]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", res.stderr)
    spills = [int(v) for v in re.findall(r"VGPRs Spill: (\d+)", res.stderr)]
    scratch = [int(v) for v in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", res.stderr)]
    assert names and len(names) == len(spills) == len(scratch)
    spilled = [(n, s) for n, s in zip(names, spills) if s]
    assert not spilled, spilled
    in_scratch = [(n, s) for n, s in zip(names, scratch) if s]
    assert not in_scratch, in_scratch
