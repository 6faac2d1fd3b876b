"""Kernel resource budget, checked at compile time (no GPU): no HIP kernel of
the library may use scratch (per-lane private memory).  Scratch on these
memory-bound kernels means spilled registers or an argument struct copied to
the stack, both of which turn register accesses into memory round trips."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpu_graphs_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src", ["bfs.hip", "msbfs.hip", "scan.hip"])
def test_no_kernel_uses_scratch(tmp_path, src):
    res = subprocess.run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", str(tmp_path / "k.o")],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", res.stderr)
    scratch = [int(x) for x in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", res.stderr)]
    assert names and len(names) == len(scratch)
    bad = [(n, s) for n, s in zip(names, scratch) if s]
    assert not bad, bad
