"""The persisted autotuner table (gpu_graphs_amd/tuning.py): entries round
trip, an entry recorded for other graph sizes is not applied, and the
committed table covers the workloads the bench and the full-size parity
tests run."""
import os

import pytest

import bench
from gpu_graphs_amd import tuning

TUNED = {"default_ms": 1.0, "best_ms": 0.5, "tried": 3, "block": 512, "edges_per_thread": 2, "alpha": 4,
         "beta": 0, "low_degree": 32, "frontier_edges": 1 << 20}


def test_record_and_lookup(tmp_path):
    path = str(tmp_path / "t.json")
    assert tuning.lookup("w", "batch", 10, 20, path) is None
    tuning.record("w", "batch", 10, 20, TUNED, path)
    tuning.record("w", "single", 10, 20, dict(TUNED, block=1024, edges_per_thread=1), path)
    assert tuning.lookup("w", "batch", 10, 20, path) == TUNED
    assert tuning.lookup("w", "single", 10, 20, path)["block"] == 1024
    assert tuning.lookup("w", "batch", 10, 21, path) is None  # another graph
    with pytest.raises(ValueError):
        tuning.lookup("w", "other", 10, 20, path)


@pytest.mark.skipif(not os.path.exists(tuning.TABLE), reason="no committed table yet")
def test_committed_table_covers_the_bench_workloads():
    t = tuning.load()
    assert "single" in t[bench.WL_RMAT22]
    assert "single" in t[bench.WL_RMAT24] and "batch" in t[bench.WL_RMAT24]
    for w in t.values():
        for e in w.values():
            assert e["n"] > 0 and e["m"] > 0 and e["tuning"]["block"] in (64, 128, 256, 512, 1024)
