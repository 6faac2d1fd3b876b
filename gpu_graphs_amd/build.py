"""Builds libgpugraphs.so (gfx950 HIP + host C++) in-tree with make.

    python -m gpu_graphs_amd.build          # library + graphBFS CLI
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, verbose=False):
    """Compiles every source (forcing a rebuild of the library objects)."""
    jobs = max(1, min(int(jobs), 16))
    cmd = ["make", "-C", HERE, f"-j{jobs}", "-B", "all"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stdout.write(res.stdout[-4000:])
        sys.stderr.write(res.stderr[-4000:])
    if res.returncode != 0:
        raise RuntimeError(f"building libgpugraphs.so failed (exit {res.returncode})")
    return os.path.join(HERE, "libgpugraphs.so")


if __name__ == "__main__":
    print(build(verbose=True))
