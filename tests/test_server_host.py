"""Host-only tests of the continuous-batching scheduler: compiles
tests/cpp/test_server.cpp against gpu_graphs_amd/csrc/server.cpp with g++
(no GPU, injected executor) and runs it, plainly and under ASan/UBSan and
TSan (host code only)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "gpu_graphs_amd", "csrc")
SRCS = [os.path.join(ROOT, "tests", "cpp", "test_server.cpp"), os.path.join(CSRC, "server.cpp"),
        os.path.join(CSRC, "errors.cpp")]


def build_and_run(tmp_path, extra):
    exe = str(tmp_path / "test_server")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I", CSRC] + extra + SRCS + ["-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAILED" not in res.stdout
    return res.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("flags", [[], ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"],
                                   ["-fsanitize=thread"]], ids=["plain", "asan_ubsan", "tsan"])
def test_scheduler(tmp_path, flags):
    out = build_and_run(tmp_path, flags)
    lines = [ln for ln in out.splitlines() if ln.strip()]
    assert len(lines) >= 10 and all(ln.split()[-1] == "ok" for ln in lines), out
