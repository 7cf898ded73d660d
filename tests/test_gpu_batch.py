"""GPU tests of batched multi-source queries, the continuous-batching server
and the autotuner.  Every lane's digest, reached count and depth must equal
the oracle's single-source answer (bit-exact); served results must equal
the same query run alone."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _expect(og, algo, src):
    val = O.bfs(og, src) if algo == "bfs" else O.sssp(og, src)
    h, r, d = O.hash_values(val, og.n + 1)
    return val, h, r, d


@pytest.fixture(scope="module")
def g14():
    og = O.rmat(14, 16, seed=21, weights=True)
    g = T.Graph.rmat(14, 16, seed=21, weights=True)
    return g, og


@pytest.mark.parametrize("algo", ["bfs", "sssp"])
@pytest.mark.parametrize("frac", [1.0, 0.25])
def test_batch_64_matches_single_source(g14, algo, frac):
    g, og = g14
    _say("case: batch", algo, "budget", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 1 << 12)
    rng = np.random.default_rng(5)
    srcs = rng.integers(0, og.n, 64).astype(np.uint32)
    srcs[3] = srcs[7]  # duplicate sources are independent lanes
    out = g.batch(algo, srcs, full=True)
    for i, s in enumerate(srcs):
        val, h, r, d = _expect(og, algo, int(s))
        assert np.array_equal(out["full"][i], val), (i, s)
        assert (out["hash"][i], out["reached"][i], out["depth"][i]) == (h, r, d), (i, s)


@pytest.mark.parametrize("weights", [True, False])
@pytest.mark.parametrize("frac", [1.0, 0.25, 0.0])
def test_batch_bfs_bottom_up_levels(weights, frac):
    """On a symmetric graph the dense levels of a 64-lane BFS run bottom-up;
    every lane's levels equal the oracle's and the top-down-only batch's,
    with the lists in HBM, split, or all read zero-copy."""
    _say("case: batch bottom-up", weights, frac)
    og = O.rmat(14, 16, seed=33, weights=weights)
    g = T.Graph.rmat(14, 16, seed=33, weights=weights)
    g.store_config(int(frac * g.m * g.edge_bytes), 1 << 10)
    rng = np.random.default_rng(6)
    srcs = rng.integers(0, og.n, 64).astype(np.uint32)
    srcs[0] = 0                        # the highest-degree vertex
    srcs[5] = srcs[9]
    deg = np.diff(og.off)
    srcs[11] = int(np.flatnonzero(deg == 0)[0])   # an isolated source
    bu = g.batch("bfs", srcs, full=True)
    assert bu["stats"].bottom_up_levels > 0
    for i, s in enumerate(srcs):
        val, h, r, d = _expect(og, "bfs", int(s))
        assert np.array_equal(bu["full"][i], val), (i, s)
        assert (bu["hash"][i], bu["reached"][i], bu["depth"][i]) == (h, r, d), (i, s)
    g.set_option("direction", 0)
    td = g.batch("bfs", srcs)
    g.set_option("direction", 1)
    assert td["stats"].bottom_up_levels == 0
    assert np.array_equal(td["hash"], bu["hash"])
    assert np.array_equal(td["depth"], bu["depth"])
    # early exit: far fewer records than the top-down batch reads
    assert bu["stats"].edges_traversed < td["stats"].edges_traversed


def test_batch_split_direction_lanes():
    """A bottom-up batch level splits its lanes: lanes whose frontier is a
    few vertices (sources on a long path hanging off an R-MAT core) run
    top-down while the core's lanes run bottom-up; every lane equals the
    oracle, across the levels where the path lanes enter the core."""
    _say("case: batch split lanes")
    core = O.rmat(16, 16, seed=41)
    n0, plen = core.n, 600
    n = n0 + plen
    deg0 = np.diff(core.off)
    hub = int(np.argmax(deg0))
    src = np.repeat(np.arange(n0), deg0.astype(np.int64))
    u = np.concatenate([src, np.arange(n0, n - 1), np.arange(n0 + 1, n), [hub, n0]])
    v = np.concatenate([core.adj.astype(np.int64), np.arange(n0 + 1, n), np.arange(n0, n - 1),
                        [n0, hub]])
    p = np.stack([u, v], 1)
    p = p[np.lexsort((p[:, 1], p[:, 0]))]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, p[:, 0] + 1, 1)
    og = O.CSR(np.cumsum(off).astype(np.uint64), p[:, 1].astype(np.uint32))
    g = T.Graph.from_arrays(og.off, og.adj)
    g.set_option("symmetric", 1)
    rng = np.random.default_rng(17)
    srcs = np.concatenate([rng.integers(0, n0, 32), n - 1 - np.arange(32) * 7]).astype(np.uint32)
    out = g.batch("bfs", srcs, full=True)
    assert out["stats"].bottom_up_levels > 0
    for i, s in enumerate(srcs):
        assert np.array_equal(out["full"][i], O.bfs(og, int(s))), (i, s)
    # the same lanes served with staggered admissions
    srv = T.Server(g, max_lanes=64, keep_full=True)
    tk = [(int(s), srv.submit("bfs", int(s))) for s in srcs[::-1]]
    srv.drain()
    for s, t in tk:
        full = np.zeros(og.n, np.uint32)
        r = srv.poll(t, full)
        assert r["status"] == 0 and np.array_equal(full, O.bfs(og, s)), s
    srv.stop()


@pytest.mark.parametrize("n", [5, 63, 1000])
def test_batch_bottom_up_small_graphs(n):
    """Bottom-up batch levels on small symmetric graphs whose vertex count is
    not a multiple of a wave (partial last group, one-word bitmaps): every
    lane equals the oracle."""
    _say("case: batch bottom-up small", n)
    rng = np.random.default_rng(n)
    k = max(1, 3 * n)
    u = rng.integers(0, n, k)
    v = rng.integers(0, n, k)
    keep = u != v
    pairs = np.unique(np.stack([np.concatenate([u[keep], v[keep]]),
                                np.concatenate([v[keep], u[keep]])], 1), axis=0)
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, pairs[:, 0] + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    og = O.CSR(off, pairs[:, 1].astype(np.uint32))
    g = T.Graph.from_arrays(og.off, og.adj)
    g.set_option("symmetric", 1)
    srcs = (np.arange(64) * 7 % n).astype(np.uint32)
    out = g.batch("bfs", srcs, full=True)
    for i, s in enumerate(srcs):
        assert np.array_equal(out["full"][i], O.bfs(og, int(s))), (n, i, s)
    # single-source direction-optimising BFS on the same graph, all edges in
    # HBM and all zero-copy
    for frac in (1.0, 0.0):
        g.store_config(int(frac * og.m * 4), 16)
        for s in srcs[:8]:
            assert np.array_equal(g.bfs(int(s))[0], O.bfs(og, int(s))), (n, frac, s)


def test_serving_bottom_up_staggered():
    """A server on a symmetric graph admits queries between levels, so its
    lanes run at different depths; levels may still run bottom-up, and each
    result equals the query alone."""
    _say("case: serving bottom-up")
    og = O.rmat(13, 16, seed=8)
    g = T.Graph.from_arrays(og.off, og.adj)
    g.set_option("symmetric", 1)
    srv = T.Server(g, max_lanes=16, keep_full=True)
    rng = np.random.default_rng(12)
    srcs = [int(s) for s in rng.integers(0, og.n, 120)]
    tk = [(s, srv.submit("bfs", s)) for s in srcs]
    srv.drain()
    for s, t in tk:
        full = np.zeros(og.n, np.uint32)
        r = srv.poll(t, full)
        assert r["status"] == 0 and np.array_equal(full, O.bfs(og, s)), s
    srv.stop()


@pytest.mark.parametrize("k", [1, 13])
def test_batch_small_k(g14, k):
    g, og = g14
    g.store_config(None, 1 << 16)
    srcs = np.arange(k, dtype=np.uint32) * 97 % og.n
    for algo in ("bfs", "sssp"):
        out = g.batch(algo, srcs)
        for i, s in enumerate(srcs):
            _, h, r, d = _expect(og, algo, int(s))
            assert (out["hash"][i], out["reached"][i], out["depth"][i]) == (h, r, d)


def test_batch_rejects_bad_input(g14):
    g, og = g14
    with pytest.raises(T.ThvError):
        g.batch("bfs", np.zeros(65, np.uint32))
    with pytest.raises(T.ThvError):
        g.batch("bfs", np.array([og.n], np.uint32))
    with pytest.raises(T.ThvError):
        g.batch("cc", np.array([0], np.uint32))


@pytest.mark.parametrize("sssp_engines", [1, 3])
def test_serving_alone_equals_mix(g14, sssp_engines):
    """Queries arrive while others run (8 lanes per engine, 300 queries, one
    or three SSSP engines): each result equals the query run alone -- on the
    oracle and through the plain single-source GPU path."""
    g, og = g14
    _say("case: serving mix, sssp engines", sssp_engines)
    g.store_config(int(0.5 * g.m * 8), 1 << 12)
    srv = T.Server(g, max_lanes=8, sssp_engines=sssp_engines)
    rng = np.random.default_rng(9)
    qs = [("bfs" if rng.random() < 0.5 else "sssp", int(rng.integers(0, og.n)))
          for _ in range(300)]
    tickets = [srv.submit(a, s) for a, s in qs]
    srv.drain()
    results = [srv.poll(t) for t in tickets]
    srv.stop()
    for (a, s), r in zip(qs, results):
        assert r["status"] == 0
        _, h, rc, d = _expect(og, a, s)
        assert (r["hash"], r["reached"], r["depth"]) == (h, rc, d), (a, s)
        assert r["submit_ms"] <= r["start_ms"] <= r["done_ms"]
    # alone on the GPU: plain single-source kernels
    for a, s in qs[:20]:
        val = g.bfs(s)[0] if a == "bfs" else g.sssp(s)[0]
        h, rc, d = O.hash_values(val, og.n + 1)
        r = results[qs.index((a, s))]
        assert (r["hash"], r["reached"], r["depth"]) == (h, rc, d)


def test_serving_keep_full_and_reuse():
    og = O.rmat(11, 8, seed=4, weights=True)
    g = T.Graph.from_arrays(og.off, og.adj, og.w)
    srv = T.Server(g, max_lanes=4, keep_full=True)
    srcs = [0, 5, 17, 99, 1000, 2047, 3, 3, 512, 64]
    tk = [(a, s, srv.submit(a, s)) for s in srcs for a in ("bfs", "sssp")]
    srv.drain()
    for a, s, t in tk:
        full = np.zeros(og.n, np.uint32)
        r = srv.poll(t, full)
        val = O.bfs(og, s) if a == "bfs" else O.sssp(og, s)
        assert r["status"] == 0 and np.array_equal(full, val), (a, s)
    srv.stop()


def test_serving_errors():
    og = O.rmat(10, 8, seed=4)
    g = T.Graph.from_arrays(og.off, og.adj)
    with pytest.raises(T.ThvError):
        T.Server(g, scratch_bytes=1024)
    srv = T.Server(g, max_lanes=2)
    with pytest.raises(T.ThvError) as e:
        srv.submit("sssp", 0)          # unweighted graph: no SSSP engine
    assert e.value.code == -7
    with pytest.raises(T.ThvError):
        srv.submit("bfs", og.n)
    with pytest.raises(T.ThvError) as e:
        g.store_config(None, 128)      # store is pinned while serving
    assert e.value.code == -6
    t = srv.submit("bfs", 1)
    srv.drain()
    assert srv.poll(t)["status"] == 0
    srv.stop()
    g.store_config(None, 128)


def test_serving_releases_results_and_pins_graph():
    """A finished query is handed over once (a second poll of its ticket is
    an error, so a long stream does not keep every result); while a server
    runs, closing the graph and autotuning are refused (THV_EBUSY) and the
    graph stays usable; after stop both work."""
    og = O.rmat(10, 8, seed=4, weights=True)
    g = T.Graph.from_arrays(og.off, og.adj, og.w)
    srv = T.Server(g, max_lanes=4, keep_full=True)
    t = srv.submit("bfs", 3)
    srv.drain()
    full = np.zeros(og.n, np.uint32)
    assert srv.poll(t, full)["status"] == 0
    assert np.array_equal(full, O.bfs(og, 3))
    with pytest.raises(T.ThvError) as e:
        srv.poll(t)
    assert e.value.code == -1
    with pytest.raises(T.ThvError) as e:
        g.close()
    assert e.value.code == -6
    with pytest.raises(T.ThvError) as e:
        g.autotune()
    assert e.value.code == -6
    t = srv.submit("sssp", 5)       # still serving after the refusals
    srv.drain()
    r = srv.poll(t)
    assert r["status"] == 0
    assert (r["hash"], r["reached"], r["depth"]) == O.hash_values(O.sssp(og, 5), og.n + 1)
    srv.stop()
    g.autotune()
    g.close()


def test_autotune_keeps_results_exact(tmp_path, g14):
    g, og = g14
    g.store_config(int(0.25 * g.m * 8), 1 << 12)
    cache = str(tmp_path / "tune.txt")
    ch = g.autotune(cache)
    assert ch["blocks_per_cu"] in (2, 4, 8) and ch["edges_per_thread"] in (4, 8)
    assert open(cache).read().count("|") == 1
    assert g.autotune(cache) == ch          # second call reads the cache
    assert open(cache).read().count("|") == 1
    assert np.array_equal(g.bfs(1)[0], O.bfs(og, 1))
    assert np.array_equal(g.sssp(1)[0], O.sssp(og, 1))
    out = g.batch("bfs", np.arange(64, dtype=np.uint32))
    for i in range(64):
        _, h, r, d = _expect(og, "bfs", i)
        assert (out["hash"][i], out["reached"][i], out["depth"][i]) == (h, r, d)
