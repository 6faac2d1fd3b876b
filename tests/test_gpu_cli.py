"""The boundary as a user of the reference meets it, on the GPU: the CLI
`graphBFS <input_file> <source_num>` (1D Partition/main.cu:10-35) and a
program written against the C++ facade with the reference's class names
(csr, gpuBFS: include/csr.h:9-23, include/gpuBFS.cuh:13-20), run on the
reference's fixture graphs.  The time line keeps the reference's format
("0 GPU BFS Time: X ms", main.cu:35) and the printed levels equal the scipy
goldens (tests/golden/fixture_levels.npz)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, fixture_path

pytestmark = pytest.mark.gpu

GRAPHBFS = os.path.join(ROOT, "gpu_graphs_amd", "graphBFS")
FACADE = os.path.join(ROOT, "gpu_graphs_amd", "facade_main")
TIME_LINE = re.compile(r"^0 GPU BFS Time: [0-9.eE+-]+ ms$", re.M)

# (fixture, extra flags, source)
CASES = [("tree", [], 1), ("tree", [], 0), ("test4", [], 0), ("test4", [], 5), ("test", ["--one-based"], 0),
         ("test", ["--one-based"], 7)]


def run(args):
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (args, r.returncode, r.stderr)
    return r.stdout


def distance_vector(stdout):
    head = "------GPU DISTANCE VECTOR------"
    assert head in stdout, stdout[:400]
    body = [line for line in stdout.split(head, 1)[1].splitlines() if line.strip()][0]
    return np.array([int(x) for x in body.replace("|", " ").split()], dtype=np.int32)


def golden(golden_levels, name, source):
    srcs = list(golden_levels[f"{name}__sources"])
    return golden_levels[f"{name}__levels"][srcs.index(source)]


@pytest.mark.parametrize("name,flags,source", CASES)
def test_graphbfs_cli(gpu, golden_levels, name, flags, source):
    path = fixture_path(name)
    out = run([GRAPHBFS, path, str(source)] + flags)
    assert TIME_LINE.search(out), out
    out = run([GRAPHBFS, path, str(source), "--print"] + flags)
    want = golden(golden_levels, name, source)
    np.testing.assert_array_equal(distance_vector(out), want)
    out = run([GRAPHBFS, path, str(source), "--json"] + flags)
    rec = json.loads(out.strip().splitlines()[-1])
    assert rec["source"] == source and rec["n"] == want.size
    assert rec["levels"] == int(want.max()) + 1
    assert rec["ms"] > 0 and rec["copy_ms"] >= 0


@pytest.mark.parametrize("name,source", [("tree", 1), ("tree", 3), ("test4", 0)])
def test_facade_program(gpu, golden_levels, name, source):
    out = run([FACADE, fixture_path(name), str(source)])
    assert TIME_LINE.search(out), out
    np.testing.assert_array_equal(distance_vector(out), golden(golden_levels, name, source))
