"""GPU parity at 2^20 vertices: the bottom-up layout of a split store (25%
budget, 2^16-edge partitions, parent hints) as the capped-budget BFS runs it,
the 64-lane BFS and SSSP batches (resident and split) and the
continuous-batching server, every result against the oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def sym20():
    og = O.rmat(20, 32, seed=7)
    g = T.Graph.rmat(20, 32, seed=7)
    return g, og


@pytest.fixture(scope="module")
def w20():
    og = O.rmat(20, 16, seed=12, weights=True)
    g = T.Graph.rmat(20, 16, seed=12, weights=True)
    return g, og


def _roots(og, k, seed):
    deg = np.diff(og.off)
    rng = np.random.default_rng(seed)
    r = rng.choice(np.flatnonzero(deg > 0), k, replace=False).astype(np.uint32)
    r[0] = 1
    return r


def test_bottom_up_layout_bfs_and_cc(sym20):
    """R-MAT 20 / ef 32 at a 25% budget with 2^16-edge partitions: the store
    takes the bottom-up layout (hints from the budget, sparsest partitions in
    HBM), BFS runs bottom-up levels, and the levels of vertex 1 and 15
    random roots equal the oracle's; CC equals label propagation."""
    g, og = sym20
    _say("case: bottom-up layout, R-MAT 20/32, budget 25%")
    g.store_config(int(0.25 * g.m * 4), 1 << 16)
    info = g.store_info()
    assert g.get_option("hint_bytes") == 4 * g.n
    assert 0 < info["device_edges"] < g.m
    roots = _roots(og, 16, 3)
    ref = O.bfs_multi(og, roots, symmetric=True, full=True)
    for i, r in enumerate(roots):
        lv, st = g.bfs(int(r))
        assert np.array_equal(lv, ref["full"][i]), (i, r)
        if i == 0:
            assert st.bottom_up_levels > 0 and st.hint_hits > 0
            assert st.host_bytes < st.topdown_host_bytes
    assert np.array_equal(g.cc()[0], O.cc_mt(og))
    # the same store read top-down only (the reference's traversal)
    g.set_option("direction", 0)
    lv, st = g.bfs(1)
    assert st.bottom_up_levels == 0
    assert np.array_equal(lv, ref["full"][0])
    g.set_option("direction", 1)


@pytest.mark.parametrize("frac", [1.0, 0.25])
def test_batch64_scale20(w20, frac):
    """64-lane BFS and SSSP batches on R-MAT 20 / ef 16 with weights, all
    lists in HBM or 25% of them: every BFS lane's value array equals the
    oracle's multi-source BFS, every SSSP lane's digest the oracle's
    delta-stepping run."""
    g, og = w20
    _say("case: batch64 R-MAT 20, budget", frac)
    g.store_config(int(frac * g.m * 8), 1 << 16)
    rng = np.random.default_rng(int(frac * 100))
    srcs = rng.integers(0, og.n, 64).astype(np.uint32)
    bfs = g.batch("bfs", srcs, full=True)
    ref = O.bfs_multi(og, srcs, symmetric=True, full=True)
    for i in range(64):
        assert np.array_equal(bfs["full"][i], ref["full"][i]), (i, srcs[i])
    assert np.array_equal(bfs["hash"], ref["hash"])
    sssp = g.batch("sssp", srcs)
    for i, s in enumerate(srcs):
        h, r, d = O.hash_values(O.sssp_delta_mt(og, int(s)), og.n + 1)
        assert (sssp["hash"][i], sssp["reached"][i], sssp["depth"][i]) == (h, r, d), (i, s)


def test_server_scale20(w20):
    """256 queries (half BFS, half SSSP, random sources) on the resident
    R-MAT 20 graph through a 64-lane server: each result equals the
    oracle's answer for the query alone."""
    g, og = w20
    _say("case: server R-MAT 20")
    g.store_config(None, 1 << 16)
    rng = np.random.default_rng(31)
    bsrc = rng.integers(0, og.n, 128).astype(np.uint32)
    ssrc = rng.integers(0, og.n, 128).astype(np.uint32)
    qs = [("bfs", int(s)) for s in bsrc] + [("sssp", int(s)) for s in ssrc]
    order = rng.permutation(len(qs))
    srv = T.Server(g, max_lanes=64)
    tk = [(qs[i], srv.submit(*qs[i])) for i in order]
    srv.drain()
    res = {q: srv.poll(t) for q, t in tk}
    srv.stop()
    exp = {}
    for i in range(0, 128, 64):
        ref = O.bfs_multi(og, bsrc[i:i + 64], symmetric=True)
        for j, s in enumerate(bsrc[i:i + 64]):
            exp[("bfs", int(s))] = (ref["hash"][j], ref["reached"][j], ref["depth"][j])
    for s in ssrc:
        exp[("sssp", int(s))] = O.hash_values(O.sssp_delta_mt(og, int(s)), og.n + 1)
    for q, r in res.items():
        assert r["status"] == 0, q
        assert (r["hash"], r["reached"], r["depth"]) == tuple(int(x) for x in exp[q]), q
