"""Writes the golden level vectors under tests/golden/ (run once, committed).

Expected levels come from scipy.sparse.csgraph.shortest_path (unweighted,
directed), an implementation independent of both this engine and its CPU
oracle.  The edge lists are parsed here with numpy, not with the engine's
loader.

Inputs:
  graphs/*.txt     the reference's own fixture graphs (graphs/ and
                   1D Partition/migration/GPU_Graphs/graphs/ of the reference)
  er_10200.txt     Erdos-Renyi workload (10,000 candidate + 200 isolated
                   vertices, 50,000 undirected edges, seed 20260002), written
                   by the engine's generator in the reference text format

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# name -> (one_based, auto_n)
FIXTURES = {
    "tree": (False, False),
    "test4": (False, False),
    "test": (True, False),
    "test2": (True, False),
    "test3": (True, False),
    "test_isolated_node": (True, False),
    "mig_tree": (True, False),
    "wiki-Vote": (True, True),
}
ER_SEED = 20260002


def parse(path, one_based, auto_n):
    with open(path) as f:
        tok = np.array(f.read().split(), dtype=np.int64)
    n = int(tok[0])
    pairs = tok[2: 2 + 2 * ((tok.size - 2) // 2)].reshape(-1, 2) - (1 if one_based else 0)
    if auto_n and pairs.size:
        n = max(n, int(pairs.max()) + 1)
    assert pairs.size == 0 or (pairs.min() >= 0 and pairs.max() < n), path
    return n, pairs


def levels_from(n, pairs, sources):
    a = csr_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    a.data[:] = 1.0  # duplicates sum; any positive weight marks an edge
    d = shortest_path(a, method="D", directed=True, unweighted=True, indices=np.asarray(sources))
    out = np.where(np.isinf(d), -1, d).astype(np.int32)
    return out.reshape(len(sources), n)


def main():
    fixtures = {}
    for name, (ob, an) in FIXTURES.items():
        n, pairs = parse(os.path.join(HERE, "graphs", f"{name}.txt"), ob, an)
        if n <= 4096:
            sources = np.arange(n, dtype=np.int32)
        else:
            outdeg = np.bincount(pairs[:, 0], minlength=n)
            rng = np.random.default_rng(7)
            cand = np.nonzero(outdeg)[0]
            sources = np.sort(rng.choice(cand, 24, replace=False)).astype(np.int32)
            sources = np.concatenate([sources, np.nonzero(outdeg == 0)[0][:8].astype(np.int32)])
        fixtures[f"{name}__sources"] = sources
        fixtures[f"{name}__levels"] = levels_from(n, pairs, sources)
        print(name, n, len(pairs), fixtures[f"{name}__levels"].shape)
    np.savez_compressed(os.path.join(HERE, "fixture_levels.npz"), **fixtures)

    sys.path.insert(0, ROOT)
    import gpu_graphs_amd as gg  # generator only; expected levels come from scipy
    g = gg.Graph.erdos_renyi(10000, 200, 50000, seed=ER_SEED)
    path = os.path.join(HERE, "er_10200.txt")
    g.write_edgelist(path)
    n, pairs = parse(path, False, False)
    lv = levels_from(n, pairs, [0])[0]
    np.save(os.path.join(HERE, "er_10200_levels.npy"), lv)
    print("er", n, len(pairs), "reached", int((lv >= 0).sum()), "levels", int(lv.max()) + 1)


if __name__ == "__main__":
    main()
