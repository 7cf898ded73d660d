"""GPU checks of the tuner (thv_tune): for BFS, SSSP, PageRank and 64-lane
batches it picks settings from its candidate sets, leaves every result
equal to the oracle's (bit-exact; PageRank within 1e-5 normalised L1 of the
FP64 fixpoint), rebuilds the store at the partition size it picks, reuses
its cache, and refuses what it cannot tune."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _valid(r, workload):
    assert r["blocks_per_cu"] in (2, 4, 8) and r["edges_per_thread"] in (4, 8)
    assert r["probe_blocks_per_cu"] in (4, 8, 16) and r["batch_blocks_per_cu"] in (4, 8, 16)
    assert r["partition_edges"] in (1 << 14, 1 << 16, 1 << 18)
    assert r["pull_degree"] in (8, 16)
    if workload in ("sssp", "batch-sssp"):
        assert r["sssp_delta"] in (8, 16, 32, 64)
    assert r["tuned_ms"] <= r["default_ms"] + 1e-9


def test_tune_bfs_split_store_with_partitions(tmp_path):
    _say("case: tune bfs, R-MAT 16 split store, partition sizes")
    og = O.rmat(16, 16, seed=4)
    g = T.Graph.rmat(16, 16, seed=4)
    g.store_config(int(0.25 * g.m * 4), 1 << 16)
    cache = str(tmp_path / "tune.txt")
    r = g.tune("bfs", partitions=True, cache_path=cache)
    _valid(r, "bfs")
    # every knob's candidates were timed: 5 shapes, 3 probe grids, 2 degrees,
    # 2 store layouts, 3 partition sizes, less the starting value of each,
    # plus the start
    assert r["candidates"] == 1 + 4 + 2 + 1 + 1 + 2 and r["from_cache"] == 0
    assert r["store_layout"] in (0, 1)
    assert g.store_info()["partitions"] > 0
    for s in (1, 5, 4097):
        assert np.array_equal(g.bfs(s)[0], O.bfs(og, s)), s
    again = g.tune("bfs", partitions=True, cache_path=cache)
    assert again["from_cache"] == 1 and again["candidates"] == 0
    assert {k: again[k] for k in again if k not in ("candidates", "from_cache")} == \
        {k: r[k] for k in r if k not in ("candidates", "from_cache")}
    assert open(cache).read().count("|") == 1


def test_tune_sssp_and_batches():
    _say("case: tune sssp and batches, weighted R-MAT 14")
    og = O.rmat(14, 16, seed=8, weights=True)
    g = T.Graph.rmat(14, 16, seed=8, weights=True)
    g.store_config(int(0.3 * g.m * 8), 1 << 14)
    r = g.tune("sssp", sources=[1, 2, 3], partitions=True, repeats=2)
    _valid(r, "sssp")
    assert r["candidates"] == 1 + 3 + 4 + 1 + 2
    assert r["store_layout"] in (0, 2)
    assert g.get_option("store_layout") == r["store_layout"]
    for s in (1, 9):
        assert np.array_equal(g.sssp(s)[0], O.sssp(og, s)), s
    srcs = np.arange(64, dtype=np.uint32) * 7
    for wl, algo in (("batch-bfs", "bfs"), ("batch-sssp", "sssp")):
        r = g.tune(wl, sources=srcs, repeats=2)
        _valid(r, wl)
        out = g.batch(algo, srcs)
        for i, s in enumerate(srcs):
            val = O.bfs(og, int(s)) if algo == "bfs" else O.sssp(og, int(s))
            assert (out["hash"][i], out["reached"][i], out["depth"][i]) == \
                O.hash_values(val, og.n + 1), (algo, s)


def test_tune_pagerank():
    _say("case: tune pagerank push and pull, directed R-MAT 14")
    og = O.rmat(14, 16, seed=2, symmetric=False)
    ref, _ = O.pr_exact(og)
    for mode, tr in (("push", False), ("pull", True)):
        g = T.Graph.rmat(14, 16, seed=2, symmetric=False, transpose=tr)
        g.store_config(int(0.25 * g.m * 4), 1 << 14)
        r = g.tune("pr", partitions=True, repeats=2, pr_iterations=5)
        _valid(r, "pr")
        x, _ = g.pr(mode)
        assert O.normalized_l1(x, ref) < PR_L1_TOL, (mode, O.normalized_l1(x, ref))


def test_tune_refusals():
    _say("case: tune refusals")
    g = T.Graph.rmat(12, 8, seed=3)
    with pytest.raises(T.ThvError) as e:
        g.tune("sssp")            # no weights
    assert e.value.code == -7
    with pytest.raises(T.ThvError) as e:
        g.tune("batch-bfs", sources=np.arange(65, dtype=np.uint32))
    assert e.value.code == -1
    with pytest.raises(T.ThvError) as e:
        g.tune("bfs", sources=[g.n])
    assert e.value.code == -1
    srv = T.Server(g, max_lanes=8)
    with pytest.raises(T.ThvError) as e:
        g.tune("bfs")
    assert e.value.code == -6
    srv.stop()
    assert g.tune("bfs", repeats=1)["candidates"] > 1
