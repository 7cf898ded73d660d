"""GPU parity of the HIP engine against the oracle (bit-exact for BFS, CC and
SSSP; PageRank within 1e-5 normalised L1 of the FP64 fixpoint).  Every call
goes through the C ABI of libthv_hip.so."""
import os

import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-5  # normalised L1 bound vs the FP64 oracle


def _say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------- fixtures ---

def test_tiny4_hand_derived(golden):
    _say("case: tiny4 fixture")
    g = T.Graph.open(os.path.join(golden, "tiny4.csr"), "csr", 4)
    assert (g.n, g.m) == (4, 6)
    assert g.bfs(1)[0].tolist() == [3, 1, 2, 5]
    assert g.cc()[0].tolist() == [0, 0, 0, 3]
    g8 = T.Graph.open(os.path.join(golden, "tiny4.csr"), "csr", 8)
    assert g8.sssp(1)[0].tolist() == [5, 1, 5, 5]
    gw = T.Graph.open(os.path.join(golden, "tiny4.wcsr"), "wcsr", 8)
    assert gw.sssp(1)[0].tolist() == [5, 1, 5, 5]
    assert gw.bfs(1)[0].tolist() == [3, 1, 2, 5]


@pytest.mark.parametrize("budget_frac", [1.0, 0.5, 0.0])
def test_rand64_all_sources(golden, budget_frac):
    _say("case: rand64 all sources, budget", budget_frac)
    og = O.read_graph(os.path.join(golden, "rand64.csr"), "csr")
    g = T.Graph.open(os.path.join(golden, "rand64.csr"), "csr", 8)
    g.store_config(int(budget_frac * g.m * 8), 16)
    assert g.store_info()["partitions"] > 1
    for s in range(og.n):
        assert np.array_equal(g.bfs(s)[0], O.bfs(og, s)), s
        assert np.array_equal(g.sssp(s)[0], O.sssp(og, s)), s
    assert np.array_equal(g.cc()[0], O.cc(og))


def test_rand64_pr(golden):
    _say("case: rand64 pagerank push/pull")
    og = O.read_graph(os.path.join(golden, "rand64.csr"), "csr")
    ref, _ = O.pr_exact(og)
    g = T.Graph.open(os.path.join(golden, "rand64.csr"), "csr", 4)
    push, _ = g.pr("push", tol=1e-7, max_iter=100)
    assert O.normalized_l1(push, ref) < PR_L1_TOL
    gc = T.Graph.open(os.path.join(golden, "rand64.csc"), "csc", 4)
    pull, _ = gc.pr("pull", tol=1e-7, max_iter=100)
    assert O.normalized_l1(pull, ref) < PR_L1_TOL


# ------------------------------------------------------------- generator ---

@pytest.mark.parametrize("scale,ef,sym,w", [(10, 8, True, True), (14, 16, True, True),
                                           (12, 16, False, False)])
def test_rmat_generator_matches_cpu(scale, ef, sym, w):
    _say("case: rmat generator", scale, ef, sym, w)
    g = T.Graph.rmat(scale, ef, seed=5, symmetric=sym, weights=w)
    off, dst, ww, _ = g.export()
    ref = O.rmat(scale, ef, seed=5, symmetric=sym, weights=w)
    assert np.array_equal(off, ref.off)
    assert np.array_equal(dst, ref.adj)
    if w:
        assert np.array_equal(ww, ref.w)


def test_rmat_transpose_matches_cpu():
    _say("case: rmat transpose")
    g = T.Graph.rmat(12, 16, seed=3, symmetric=False, transpose=True)
    off, src, _, od = g.export()
    ref = O.transpose(O.rmat(12, 16, seed=3, symmetric=False))
    assert np.array_equal(off, ref.off)
    assert np.array_equal(src, ref.adj)
    assert np.array_equal(od, ref.outdeg)


# ---------------------------------------------------------- split store ---

@pytest.fixture(scope="module")
def rmat16():
    g = T.Graph.rmat(16, 16, seed=1, weights=True)
    og = O.rmat(16, 16, seed=1, weights=True)
    return g, og


def test_rmat20_plain_bfs_from_vertex_1():
    """The plain configuration: BFS from vertex 1 on the undirected R-MAT
    scale 20 / edge factor 16 graph, every edge in HBM; levels bit-exact,
    with and without bottom-up levels; CC of the same graph."""
    _say("case: rmat20 plain bfs")
    g = T.Graph.rmat(20, 16, seed=1)
    og = O.rmat(20, 16, seed=1)
    g.store_config(None, 1 << 16)
    ref = O.bfs(og, 1)
    lev, st = g.bfs(1)
    assert np.array_equal(lev, ref)
    assert st.bottom_up_levels > 0
    g.set_option("direction", 0)
    lev_td, _ = g.bfs(1)
    g.set_option("direction", 1)
    assert np.array_equal(lev_td, ref)
    lab, _ = g.cc()
    assert np.array_equal(lab, O.cc(og))


@pytest.mark.parametrize("frac,part", [(1.0, 1 << 16), (0.25, 1 << 16), (0.25, 1024),
                                       (0.0, 4096)])
def test_bfs_cc_sssp_split_store(rmat16, frac, part):
    g, og = rmat16
    _say("case: rmat16 split store", frac, part)
    g.store_config(int(frac * g.m * g.edge_bytes), part)
    info = g.store_info()
    if frac == 0.0:
        assert info["device_edges"] == 0
    for s in (1, 7, 12345):
        lv, st = g.bfs(s)
        ref = O.bfs(og, s)
        assert np.array_equal(lv, ref), s
        reached = ref != og.n + 1
        assert st.edges_traversed == int(og.degrees()[reached].sum())
        if frac == 1.0:
            assert st.host_bytes == 0
        d, _ = g.sssp(s)
        assert np.array_equal(d, O.sssp(og, s)), s
    assert np.array_equal(g.cc()[0], O.cc(og))


@pytest.mark.parametrize("frac", [1.0, 0.25, 0.0])
def test_bfs_direction_optimizing(rmat16, frac):
    """Bottom-up levels on a symmetric graph give the same levels; edges
    examined drop below the traversed count."""
    g, og = rmat16
    _say("case: direction-optimising bfs", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    assert g.get_option("symmetric") == 1.0 and g.get_option("direction") == 1.0
    deg = og.degrees()
    for s in (1, 7, 12345, 65535):
        ref = O.bfs(og, s)
        lv, st = g.bfs(s)
        assert np.array_equal(lv, ref), s
        assert st.edges_traversed == int(deg[ref != og.n + 1].sum())
        assert st.edges_examined <= st.edges_traversed
        if s == 1:
            assert st.bottom_up_levels > 0
            assert st.edges_examined < st.edges_traversed
    g.set_option("direction", 0)
    lv, st = g.bfs(1)
    assert st.bottom_up_levels == 0 and np.array_equal(lv, O.bfs(og, 1))
    g.set_option("do_alpha", 2.0)   # switch earlier: still exact
    g.set_option("direction", 1)
    assert np.array_equal(g.bfs(7)[0], O.bfs(og, 7))
    g.set_option("do_alpha", 14.0)


@pytest.mark.parametrize("frac", [1.0, 0.25, 0.0])
def test_bfs_frontier_summary(rmat16, frac):
    """The bottom-up scans with a frontier summary at every level (do_summary
    1e-9), and never (0): the oracle's levels and the same counts either
    way; on the beads graph, whose bottom-up frontiers are one clique."""
    g, og = rmat16
    _say("case: frontier summary", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    try:
        bu = 0
        for s in (1, 12345, 65535):
            ref = O.bfs(og, s)
            g.set_option("do_summary", 1e-9)
            lv1, st1 = g.bfs(s)
            g.set_option("do_summary", 0)
            lv0, st0 = g.bfs(s)
            assert np.array_equal(lv1, ref) and np.array_equal(lv0, ref), s
            assert (st1.edges_traversed, st1.iterations, st1.bottom_up_levels) == \
                (st0.edges_traversed, st0.iterations, st0.bottom_up_levels), s
            bu += st1.bottom_up_levels
        assert bu > 0
    finally:
        g.set_option("do_summary", 1024)
    ob = _beads(40, 24)
    gb = T.Graph.from_arrays(ob.off, ob.adj)
    gb.set_option("symmetric", 1)
    gb.set_option("do_alpha", 1000.0)
    gb.set_option("do_beta", 1.0)
    gb.set_option("do_summary", 1e-9)
    gb.store_config(int(frac * ob.m * 4), 64)
    for s in (0, 500):
        assert np.array_equal(gb.bfs(s)[0], O.bfs(ob, s)), s


@pytest.mark.parametrize("frac", [1.0, 0.25])
def test_bfs_kernel_timing_off(rmat16, frac):
    """Without HIP events (kernel_timing 0) a direction-optimising BFS gives
    the levels and every count of a timed run; only the kernel times are 0.
    Its top-down levels' host partitions come from the counters the next
    level's count pass publishes: a split store reports some."""
    g, og = rmat16
    _say("case: bfs without kernel timing", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    times = ("ms", "expand_ms", "probe_ms", "bottom_up_ms")
    try:
        for s in (1, 12345):
            ref = O.bfs(og, s)
            g.set_option("kernel_timing", 1)
            lv1, st1 = g.bfs(s)
            g.set_option("kernel_timing", 0)
            assert g.get_option("kernel_timing") == 0.0
            lv0, st0 = g.bfs(s)
            assert np.array_equal(lv1, ref) and np.array_equal(lv0, ref), s
            d1, d0 = st1.as_dict(), st0.as_dict()
            assert {k: d0[k] for k in d0 if k not in times} == \
                {k: d1[k] for k in d1 if k not in times}, s
            assert st0.expand_ms == st0.probe_ms == st0.bottom_up_ms == 0.0
            assert st1.expand_ms > 0 and st1.bottom_up_levels > 0 and st1.bottom_up_ms > 0
            if frac < 1.0:
                assert st1.host_partitions_touched > 0, s
    finally:
        g.set_option("kernel_timing", 1)


@pytest.mark.parametrize("frac", [1.0, 0.25, 0.0])
def test_bfs_values_after_other_runs(rmat16, frac):
    """The value array is shared by every run on a graph: levels from a
    source in a small component right after a BFS over the giant one, and
    after CC (labels in the same array), are exact, top-down and
    direction-optimising."""
    g, og = rmat16
    _say("case: bfs values after other runs", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    lab = O.cc(og)
    sizes = np.bincount(lab, minlength=og.n)
    deg = og.degrees()
    small = np.flatnonzero((sizes[lab] >= 2) & (sizes[lab] <= 8) & (deg > 0))
    assert small.size > 0
    v = int(small[0])
    try:
        for direction in (1, 0):
            g.set_option("direction", direction)
            for s in (1, v, 1, v):
                assert np.array_equal(g.bfs(s)[0], O.bfs(og, s)), (direction, s)
            assert np.array_equal(g.cc()[0], lab)
            assert np.array_equal(g.bfs(v)[0], O.bfs(og, v)), (direction, v)
            _, st = g.bfs(1)
            assert np.array_equal(g.bfs(1)[0], O.bfs(og, 1))
            if direction:
                assert st.bottom_up_levels > 0
    finally:
        g.set_option("direction", 1)


@pytest.mark.parametrize("layout", [1, 2])
def test_store_layouts_exact(layout):
    """Both residency orders of a split store -- densest partitions first,
    or sparsest first with the short host lists copied to HBM -- give the
    oracle's SSSP distances, BFS levels and CC labels on 8-byte records, and
    BFS levels on the symmetric 4-byte graph with parent hints."""
    _say("case: store layout", layout)
    og = O.rmat(15, 16, seed=3, weights=True)
    g = T.Graph.rmat(15, 16, seed=3, weights=True)
    g.store_config(int(0.3 * g.m * 8), 1 << 12)
    auto = g.store_info()
    g.set_option("store_layout", layout)
    assert g.get_option("store_layout") == layout
    info = g.store_info()
    assert 0 < info["device_edges"] < g.m
    # 8-byte records choose densest first by themselves
    assert (info == auto) == (layout == 1)
    for s in (1, 77, 4321):
        assert np.array_equal(g.sssp(s)[0], O.sssp(og, s)), s
        assert np.array_equal(g.bfs(s)[0], O.bfs(og, s)), s
    assert np.array_equal(g.cc()[0], O.cc(og))
    og4 = O.rmat(15, 16, seed=3)
    g4 = T.Graph.rmat(15, 16, seed=3)
    g4.store_config(int(0.25 * g4.m * 4), 1 << 12)
    g4.set_option("store_layout", layout)
    assert g4.get_option("hint_bytes") > 0
    for s in (1, 77, 4321):
        assert np.array_equal(g4.bfs(s)[0], O.bfs(og4, s)), s


def _beads(k, s):
    """k cliques of s vertices in a row, clique b's last vertex joined to
    clique b+1's first: from vertex 0 the frontier alternates between one
    vertex and a whole clique."""
    pairs = []
    for b in range(k):
        c = np.arange(b * s, (b + 1) * s)
        u, v = np.meshgrid(c, c)
        keep = u != v
        pairs.append(np.stack([u[keep], v[keep]], 1))
        if b + 1 < k:
            pairs.append(np.array([[b * s + s - 1, (b + 1) * s], [(b + 1) * s, b * s + s - 1]]))
    p = np.concatenate(pairs)
    p = p[np.lexsort((p[:, 1], p[:, 0]))]
    off = np.zeros(k * s + 1, np.uint64)
    np.add.at(off, p[:, 0] + 1, 1)
    return O.CSR(np.cumsum(off).astype(np.uint64), p[:, 1].astype(np.uint32))


@pytest.mark.parametrize("frac", [1.0, 0.5])
def test_bfs_beads_alternating_directions(frac):
    """Levels alternate bottom-up (a clique's worth of frontier) and
    top-down (one bridge vertex), so every probe after the first reads the
    candidates an earlier probe left, across top-down levels in between:
    the levels equal the oracle's."""
    og = _beads(60, 24)
    g = T.Graph.from_arrays(og.off, og.adj)
    g.set_option("symmetric", 1)
    g.set_option("do_alpha", 1000.0)   # bottom-up as soon as the frontier grows
    g.set_option("do_beta", 1.0)       # top-down as soon as it shrinks
    g.set_option("do_link_gbps", 0.001)
    g.store_config(int(frac * og.m * 4), 64)
    _say("case: beads", frac)
    for s in (0, 700, og.n - 1):
        lv, st = g.bfs(s)
        assert np.array_equal(lv, O.bfs(og, s)), s
        if frac == 1.0 and s == 0:
            assert st.bottom_up_levels >= 20, st.bottom_up_levels
            assert st.iterations - st.bottom_up_levels >= 20, st.iterations


@pytest.fixture(scope="module")
def rmat16u():
    """Symmetric R-MAT 16 with 4-byte records: the bottom-up layout (parent
    hints, sparsest partitions first) applies to it."""
    return T.Graph.rmat(16, 16, seed=1), O.rmat(16, 16, seed=1)


@pytest.mark.parametrize("frac", [1.0, 0.25, 0.0])
@pytest.mark.parametrize("pull_degree", [0, 1, 5, 16])
@pytest.mark.parametrize("records", [4, 8])
def test_bfs_bottom_up_full_scan_split(rmat16, rmat16u, frac, pull_degree, records):
    """Bottom-up levels split between the full-list scan (degree <= D) and
    the early-exit scan give the same levels for every split point; a slow
    link rate makes the split-store rule choose bottom-up at this size."""
    g, og = rmat16u if records == 4 else rmat16
    _say("case: hybrid bottom-up", frac, pull_degree)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    g.set_option("do_pull_degree", pull_degree)
    g.set_option("do_link_gbps", 0.5)
    deg = og.degrees()
    try:
        for s in (1, 7, 65535):
            ref = O.bfs(og, s)
            lv, st = g.bfs(s)
            assert np.array_equal(lv, ref), s
            assert st.edges_traversed == int(deg[ref != og.n + 1].sum())
            if s == 1:
                assert st.bottom_up_levels > 0
            if frac == 1.0:
                assert st.host_bytes == 0
    finally:
        g.set_option("do_pull_degree", 16)
        g.set_option("do_link_gbps", 55)
    assert g.get_option("do_pull_degree") == 16.0


def test_parent_hints(rmat16, rmat16u):
    """Parent hints take n x 4 bytes of the budget on a split store of a
    symmetric graph with 4-byte records, place vertices without reading
    their lists, and leave the levels unchanged; switching them off gives
    the bytes back to the partitions.  A split store with 8-byte records
    gets no hints; a resident store keeps them beside the lists, for either
    record size."""
    gw, ogw = rmat16
    gw.store_config(int(0.25 * gw.m * gw.edge_bytes), 4096)
    assert gw.get_option("hint_bytes") == 0
    gw.store_config(None, 4096)
    assert gw.get_option("hint_bytes") == 4 * gw.n
    assert gw.store_info()["device_edges"] == gw.m
    for s in (1, 7):
        lv, st = gw.bfs(s)
        assert np.array_equal(lv, O.bfs(ogw, s)), s
        if s == 1:
            assert st.bottom_up_levels > 0 and st.hint_hits > 0
    g, og = rmat16u
    _say("case: parent hints")
    budget = int(0.25 * g.m * g.edge_bytes)
    g.store_config(budget, 4096)
    assert g.get_option("parent_hints") == 1.0
    assert g.get_option("hint_bytes") == 4 * g.n
    dev_with = g.store_info()["device_edges"]
    assert dev_with * g.edge_bytes <= budget - 4 * g.n
    g.set_option("do_link_gbps", 0.5)   # bottom-up levels at this size
    try:
        for s in (1, 7, 65535):
            lv, st = g.bfs(s)
            assert np.array_equal(lv, O.bfs(og, s)), s
            if s == 1:
                assert st.bottom_up_levels > 0 and st.hint_hits > 0
                assert st.probe_launches >= st.bottom_up_levels
                assert st.probe_hbm_bytes > 0 and st.probe_ms > 0
        g.set_option("parent_hints", 0)   # rebuilds the store without hints
        assert g.get_option("hint_bytes") == 0
        assert g.store_info()["device_edges"] > dev_with
        lv, st = g.bfs(1)
        assert np.array_equal(lv, O.bfs(og, 1)) and st.hint_hits == 0
        g.set_option("parent_hints", 1)
        assert g.get_option("hint_bytes") == 4 * g.n
        g.set_option("direction", 0)      # no bottom-up levels: no hints
        assert g.get_option("hint_bytes") == 0
        g.set_option("direction", 1)
        assert g.get_option("hint_bytes") == 4 * g.n
        assert np.array_equal(g.bfs(12345)[0], O.bfs(og, 12345))
    finally:
        g.set_option("do_link_gbps", 55)
    with pytest.raises(T.ThvError):
        g.set_option("hint_bytes", 1)
    with pytest.raises(T.ThvError):
        g.set_option("do_pull_degree", 17)
    g.store_config(int(0.0 * g.m), 4096)   # budget below 8n bytes: no hints
    assert g.get_option("hint_bytes") == 0
    gd = T.Graph.rmat(12, 8, symmetric=False)
    gd.store_config(int(0.25 * gd.m * 4), 512)
    assert gd.get_option("hint_bytes") == 0


def test_side_area_near_minimum_hint_budget():
    """A bottom-up layout budget just above the hints' minimum (8n bytes):
    the short-list side area cannot fit even with every partition evicted,
    so the store keeps its partitions and goes without the side area (it
    used to evict them all); BFS stays exact."""
    _say("case: side area at the minimum hint budget")
    og = O.rmat(16, 16, seed=1)
    g = T.Graph.from_arrays(og.off, og.adj)
    g.set_option("symmetric", 1)
    budget = 8 * og.n + (64 << 10)
    deg = np.diff(og.off).astype(np.int64)
    short = int(deg[(deg >= 1) & (deg <= 16)].sum())
    assert short * 4 > budget - 4 * og.n     # the side area alone exceeds the room
    g.store_config(budget, 1024)
    assert g.get_option("hint_bytes") == 4 * og.n
    info = g.store_info()
    assert info["device_edges"] > 0
    assert info["device_edges"] * 4 <= budget - 4 * og.n
    for s in (1, 7, 999):
        assert np.array_equal(g.bfs(s)[0], O.bfs(og, s)), s


def test_symmetric_hint_on_arrays():
    og = O.rmat(12, 16, seed=11)
    g = T.Graph.from_arrays(og.off, og.adj)
    assert g.get_option("symmetric") == 0.0
    g.set_option("symmetric", 1)
    g.store_config(int(0.2 * g.m * 4), 512)
    for s in (0, 1, 100):
        lv, st = g.bfs(s)
        assert np.array_equal(lv, O.bfs(og, s))
    with pytest.raises(T.ThvError):
        g.set_option("do_beta", -1)
    gc = T.Graph.rmat(10, 8, symmetric=False, transpose=True)
    with pytest.raises(T.ThvError):
        gc.set_option("symmetric", 1)


@pytest.mark.parametrize("frac", [1.0, 0.25])
def test_sssp_delta_stepping(rmat16, frac):
    """Delta-stepping (any bucket width) reaches the frontier Bellman-Ford
    fixpoint: the distances equal the oracle's Dijkstra bit for bit; wider
    buckets relax more edges, width 0 is the reference's order."""
    g, og = rmat16
    _say("case: sssp delta-stepping", frac)
    g.store_config(int(frac * g.m * g.edge_bytes), 4096)
    try:
        for s in (1, 7, 65535):
            ref = O.sssp(og, s)
            edges = {}
            for delta in (0, 1, 7, 32, 1000, 1 << 31):
                g.set_option("sssp_delta", delta)
                d, st = g.sssp(s)
                assert np.array_equal(d, ref), (s, delta)
                edges[delta] = st.edges_traversed
            # every bucket width relaxes each reached vertex at least once
            reached = ref != og.n + 1
            assert edges[1] >= int(og.degrees()[reached].sum())
    finally:
        g.set_option("sssp_delta", 16)
    assert g.get_option("sssp_delta") == 16.0


@pytest.mark.parametrize("scale,ef,frac", [(16, 16, 1.0), (16, 16, 0.25), (14, 1, 0.5),
                                           (12, 2, 0.0), (14, 4, 0.6), (15, 8, 0.3)])
def test_cc_union_find(scale, ef, frac):
    """Union-find CC of a symmetric graph gives the component minima, the
    labels of min-label propagation: equal to the oracle and to the
    propagation path, on a giant-component R-MAT and on sparse ones with
    many components; with parent hints (split 4-byte stores of at least 8n
    bytes) a host list's vertex links its hint first."""
    _say("case: cc union-find", scale, ef, frac)
    g = T.Graph.rmat(scale, ef, seed=5, symmetric=True)
    og = O.rmat(scale, ef, seed=5)
    g.store_config(int(frac * g.m * g.edge_bytes), 1024)
    if (scale, ef) in ((16, 16), (14, 4), (15, 8)) and frac < 1.0:
        assert g.get_option("hint_bytes") > 0
    ref = O.cc(og)
    assert g.get_option("cc_union") == 1.0
    lab, st = g.cc()
    assert np.array_equal(lab, ref)
    assert st.edges_traversed <= g.m
    g.set_option("cc_union", 0)
    lab2, st2 = g.cc()
    assert np.array_equal(lab2, ref)
    g.set_option("cc_union", 1)
    assert len(np.unique(ref)) > 1


def test_cc_union_find_step_guard_fallback():
    """A union-find link that needs more steps than the guard allows raises
    the device flag; CC then reruns by label propagation and the labels
    are still the component minima.  The guard is lowered from the host,
    so every kernel still ends normally."""
    _say("case: cc union-find step guard")
    g = T.Graph.rmat(14, 16, seed=9, symmetric=True)
    og = O.rmat(14, 16, seed=9)
    g.store_config(None, 1 << 12)
    ref = O.cc(og)
    lab, st = g.cc()
    assert np.array_equal(lab, ref)
    assert st.iterations <= 2            # union-find: two passes
    g.set_option("cc_uf_steps", 1)
    assert g.get_option("cc_uf_steps") == 1.0
    lab1, st1 = g.cc()
    assert np.array_equal(lab1, ref)
    assert st1.iterations > 2            # the label-propagation rerun
    with pytest.raises(T.ThvError):
        g.set_option("cc_uf_steps", 0)
    g.set_option("cc_uf_steps", 1 << 20)
    lab2, st2 = g.cc()
    assert np.array_equal(lab2, ref) and st2.iterations <= 2


def test_pr_push_pull_rmat():
    _say("case: rmat14 directed pagerank")
    og = O.rmat(14, 16, seed=2, symmetric=False)
    ref, _ = O.pr_exact(og)
    g = T.Graph.rmat(14, 16, seed=2, symmetric=False)
    g.store_config(int(0.25 * g.m * 4), 1 << 12)
    push, st = g.pr("push")
    assert O.normalized_l1(push, ref) < PR_L1_TOL, O.normalized_l1(push, ref)
    gc = T.Graph.rmat(14, 16, seed=2, symmetric=False, transpose=True)
    gc.store_config(int(0.25 * gc.m * 4), 1 << 12)
    pull, st2 = gc.pr("pull")
    assert O.normalized_l1(pull, ref) < PR_L1_TOL, O.normalized_l1(pull, ref)


# ----------------------------------------------------------- edge cases ---

def test_source_out_of_range(golden):
    _say("case: bad source")
    g = T.Graph.open(os.path.join(golden, "tiny4.csr"), "csr", 4)
    with pytest.raises(T.ThvError) as e:
        g.bfs(4)
    assert e.value.code == -1


def test_sssp_needs_weights(golden):
    g = T.Graph.open(os.path.join(golden, "tiny4.csr"), "csr", 4)
    with pytest.raises(T.ThvError) as e:
        g.sssp(1)
    assert e.value.code == -7


def test_single_vertex_and_isolated():
    _say("case: single vertex / isolated")
    g = T.Graph.from_arrays(np.array([0, 0], np.uint64), np.zeros(0, np.uint32))
    assert g.bfs(0)[0].tolist() == [1]
    assert g.cc()[0].tolist() == [0]
    off = np.array([0, 1, 1, 1, 2], np.uint64)
    dst = np.array([3, 0], np.uint32)
    g = T.Graph.from_arrays(off, dst)
    assert g.bfs(0)[0].tolist() == [1, 5, 5, 2]
    assert g.cc()[0].tolist() == [0, 1, 2, 0]
    # symmetric (direction-optimising path): an isolated source ends at
    # level 0; the values of the call before must not leak into the next
    g.set_option("symmetric", 1)
    assert g.bfs(0)[0].tolist() == [1, 5, 5, 2]
    assert g.bfs(1)[0].tolist() == [5, 1, 5, 5]
    assert g.bfs(3)[0].tolist() == [2, 5, 5, 1]
    assert g.cc()[0].tolist() == [0, 1, 2, 0]
    assert g.bfs(2)[0].tolist() == [5, 5, 1, 5]


def test_hub_spanning_many_chunks():
    """A star whose hub list spans many expansion chunks, plus a long path
    (many levels)."""
    _say("case: star + path")
    n = 22000
    # hub 0 -> 1..19999, path 19999 -> 20000 -> ... -> 21999
    adj = [list(range(1, 20000))] + [[] for _ in range(n - 1)]
    adj[19999].append(20000)
    for v in range(20000, n - 1):
        adj[v].append(v + 1)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(a) for a in adj])
    dst = np.array([x for a in adj for x in a], np.uint32)
    og = O.CSR(off, dst)
    g = T.Graph.from_arrays(off, dst)
    for frac in (1.0, 0.0):
        g.store_config(int(frac * g.m * 4), 512)
        assert np.array_equal(g.bfs(0)[0], O.bfs(og, 0))
        assert np.array_equal(g.cc()[0], O.cc(og))


def test_symmetric_star_path_small_frontiers():
    """Direction-optimising BFS on a symmetric star (hub 0, 20,000 leaves)
    with a 2,000-vertex path hanging off one leaf: from a leaf, level 0 is a
    one-list expansion, level 1's frontier is the hub alone (a listed queue
    whose one list spans many expansion chunks), then the leaves, then one
    path vertex per level (listed queues); from the hub and from the path's
    end as well, all edges in HBM and all on the host."""
    _say("case: symmetric star + path")
    n = 22000
    adj = [[] for _ in range(n)]
    for v in range(1, 20000):
        adj[0].append(v)
        adj[v].append(0)
    for v in range(19999, n - 1):
        adj[v].append(v + 1)
        adj[v + 1].append(v)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(sorted(a)) for a in adj])
    dst = np.array([x for a in adj for x in sorted(a)], np.uint32)
    og = O.CSR(off, dst)
    g = T.Graph.from_arrays(off, dst)
    g.set_option("symmetric", 1)
    for frac in (1.0, 0.0):
        g.store_config(int(frac * g.m * 4), 512)
        for s in (5, 0, n - 1, 19999):
            lv, st = g.bfs(s)
            assert np.array_equal(lv, O.bfs(og, s)), (frac, s)
            assert st.edges_traversed == int(og.degrees()[lv != n + 1].sum()), (frac, s)


@pytest.mark.parametrize("k", [4096, 4097])
def test_bfs_listed_queue_limit(k):
    """Top-down levels of exactly 4,096 frontier vertices (listed from the
    bitmap in one pass) and of 4,097 (the tile-table compaction): a hub 0
    with k spokes, each spoke with one more neighbour.  With every list and
    the parent hints in HBM (the default store) the edge-ratio rule keeps
    the spokes' level top-down (do_alpha tiny; only a level that finds
    every edge already visited may run bottom-up); with every list on the
    host the link-cost rule chooses, and the levels must match either way."""
    _say("case: listed queue limit", k)
    n = 2 * k + 1
    adj = [[] for _ in range(n)]
    for i in range(1, k + 1):
        adj[0].append(i)
        adj[i] += [0, k + i]
        adj[k + i].append(i)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(a) for a in adj])
    dst = np.array([x for a in adj for x in sorted(a)], np.uint32)
    og = O.CSR(off, dst)
    g = T.Graph.from_arrays(off, dst)
    g.set_option("symmetric", 1)
    g.set_option("do_alpha", 1e-9)
    try:
        for budget in (None, 0):
            g.store_config(budget, 256)
            assert (g.store_info()["device_edges"] == g.m) == (budget is None)
            for s in (0, 1, n - 1):
                lv, st = g.bfs(s)
                assert np.array_equal(lv, O.bfs(og, s)), (budget, s)
                if budget is None:
                    assert st.bottom_up_levels <= 1, s
                assert st.edges_traversed == int(og.degrees()[lv != n + 1].sum()), (budget, s)
    finally:
        g.set_option("do_alpha", 14.0)


def test_directed_cc_is_min_ancestor():
    _say("case: directed cc")
    og = O.rmat(12, 8, seed=9, symmetric=False)
    g = T.Graph.from_arrays(og.off, og.adj)
    g.store_config(int(0.3 * g.m * 4), 256)
    assert np.array_equal(g.cc()[0], O.cc_minreach(og))
