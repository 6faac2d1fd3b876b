"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed), errors come back as status codes with messages, and the
CLI surface `graphBFS <input_file> <source_num>` parses its arguments; bench.py
refuses a run on more than one GPU."""
import ctypes as C
import glob
import os
import re
import subprocess
import sys

import pytest

import gpu_graphs_amd as gg
from conftest import ROOT
from gpu_graphs_amd import _ffi


def declared_functions():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        with open(hdr) as f:
            text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(gb_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_abi():
    names = declared_functions()
    assert "gb_bfs" in names and "gb_graph_load_edgelist" in names and len(names) >= 25


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_ffi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(_ffi.SIGNATURES)


def test_version_and_errors():
    assert b"gfx950" in gg.lib().gb_version()
    with pytest.raises(gg.GBError) as e:
        gg.Graph.from_file("/nonexistent/graph.txt")
    assert e.value.code == _ffi.GB_EIO and "unable to open" in str(e.value)
    h = C.c_void_p()
    assert gg.lib().gb_graph_info(None, None, None) == _ffi.GB_EINVAL
    assert b"null" in gg.lib().gb_last_error()
    assert gg.lib().gb_bfs(None, 0, None, None, None) == _ffi.GB_EINVAL
    assert gg.lib().gb_graph_generate_rmat(0, 16, .57, .19, .19, 1, 1, C.byref(h)) == _ffi.GB_EINVAL


def test_device_count_without_gpu_is_not_an_error():
    assert gg.device_count() >= 0


def test_cli_usage():
    exe = os.path.join(ROOT, "gpu_graphs_amd", "graphBFS")
    assert os.path.exists(exe), "graphBFS not built"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([exe, "/nonexistent.txt", "0"], capture_output=True, text=True)
    assert r.returncode == 2 and "unable to open" in r.stderr


def test_facade_program_usage():
    exe = os.path.join(ROOT, "gpu_graphs_amd", "facade_main")
    assert os.path.exists(exe), "facade_main not built"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    # The bench measures one GPU; asked for more (or started as one rank of
    # several) it exits with status 2 and says why, before touching a device.
    bench = os.path.join(ROOT, "bench.py")
    r = subprocess.run([sys.executable, bench, "--gpus", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU only" in r.stderr
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, bench], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 2 and "one GPU only" in r.stderr
