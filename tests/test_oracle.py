"""Pins the oracle (CPU, no GPU): the reference converter's fixture bytes and
their hand-derived answers, scipy.sparse.csgraph routines, brute force on
tiny graphs, a sparse linear solve for PageRank, known-answer vectors for
the hash, and agreement between independent algorithms."""
import itertools
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as spla

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert")


# ------------------------------------------------------------ formats ---

def test_fixture_bytes_read(golden):
    g = O.read_graph(os.path.join(golden, "tiny4.csr"), "csr")
    assert g.off.tolist() == [0, 2, 4, 5, 6]
    assert g.adj.tolist() == [1, 2, 2, 2, 0, 3]
    assert g.w.tolist() == [20] * 6
    c = O.read_graph(os.path.join(golden, "tiny4.csc"), "csc")
    assert c.off.tolist() == [0, 1, 2, 5, 6]
    assert c.adj.tolist() == [2, 0, 0, 1, 1, 3]
    assert c.outdeg.tolist() == [2, 2, 1, 1]
    w = O.read_graph(os.path.join(golden, "tiny4.wcsr"), "wcsr")
    assert w.adj.tolist() == g.adj.tolist() and w.w.tolist() == [20] * 6


@pytest.mark.parametrize("name", ["tiny4", "rand64"])
@pytest.mark.parametrize("fmt", ["csr", "csc", "wcsr"])
def test_writer_reproduces_converter_bytes(golden, tmp_path, name, fmt):
    src = os.path.join(golden, f"{name}.{fmt}")
    g = O.read_graph(src, fmt)
    out = str(tmp_path / f"x.{fmt}")
    O.write_graph(out, fmt, g)
    assert open(out, "rb").read() == open(src, "rb").read()


def _text_to_csr(edges, n):
    """CSR of a text edge list as the converter builds it: sorted by
    (src, dst), duplicates and self-loops kept."""
    e = sorted(edges)
    off = np.zeros(n + 1, np.uint64)
    for u, _ in e:
        off[u + 1] += 1
    off = np.cumsum(off).astype(np.uint64)
    return off, np.array([v for _, v in e], np.uint32)


def test_reader_matches_text_edge_list(golden):
    lines = open(os.path.join(golden, "rand64.txt")).read().split("\n")[1:]
    edges = [tuple(map(int, l.split())) for l in lines if l.strip()]
    n = max(max(u, v) for u, v in edges) + 1
    off, adj = _text_to_csr(edges, n)
    g = O.read_graph(os.path.join(golden, "rand64.csr"), "csr")
    assert np.array_equal(g.off, off) and np.array_equal(g.adj, adj)
    c = O.read_graph(os.path.join(golden, "rand64.csc"), "csc")
    toff, tadj = _text_to_csr([(v, u) for u, v in edges], n)
    assert np.array_equal(c.off, toff) and np.array_equal(c.adj, tadj)
    assert np.array_equal(c.outdeg, np.diff(off).astype(np.uint32))


@pytest.mark.skipif(not os.path.exists(REF_CONVERT), reason="reference converter not built")
def test_reference_converter_on_fresh_graph(tmp_path):
    rng = np.random.default_rng(11)
    edges = [(int(u), int(v)) for u, v in rng.integers(0, 200, size=(1500, 2))]
    edges.append((199, 0))
    txt = tmp_path / "g.txt"
    txt.write_text("# header\n" + "\n".join(f"{u} {v}" for u, v in edges) + "\n")
    subprocess.run([REF_CONVERT, "--input", str(txt), "--skip", "1", "--type", "csr"],
                   check=True, capture_output=True)
    g = O.read_graph(str(tmp_path / "g.csr"), "csr")
    off, adj = _text_to_csr(edges, 200)
    assert np.array_equal(g.off, off) and np.array_equal(g.adj, adj)
    # the transpose of our reader's CSR equals the converter's CSC
    subprocess.run([REF_CONVERT, "--input", str(txt), "--skip", "1", "--type", "csc"],
                   check=True, capture_output=True)
    c = O.read_graph(str(tmp_path / "g.csc"), "csc")
    t = O.transpose(g)
    assert np.array_equal(c.off, t.off) and np.array_equal(c.adj, t.adj)
    assert np.array_equal(c.outdeg, t.outdeg)


# ------------------------------------------------- hand-derived answers ---

def test_tiny4_answers(golden):
    g = O.read_graph(os.path.join(golden, "tiny4.csr"), "csr")
    assert O.bfs(g, 1).tolist() == [3, 1, 2, 5]
    assert O.cc(g).tolist() == [0, 0, 0, 3]
    assert O.cc_minreach(g).tolist() == [0, 0, 0, 3]
    assert O.sssp(g, 1).tolist() == [5, 1, 5, 5]
    assert O.sssp_bf(g, 1).tolist() == [5, 1, 5, 5]


def test_mix64_known_answer():
    # SplitMix64 seeded with 0: first output 0xE220A8397B1DCDAF, second
    # 0x6E789E6AA1B965F4 (the generator adds the golden gamma before mixing).
    assert O.mix64(0) == 0xE220A8397B1DCDAF
    assert O.mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    z = np.array([0, 0x9E3779B97F4A7C15, 12345], np.uint64)
    assert O.mix64_np(z).tolist() == [O.mix64(int(x)) for x in z]


def test_hash_values_by_hand():
    val = np.array([1, 5, 3, 5], np.uint32)
    h, r, d = O.hash_values(val, 5)
    exp = (O.mix64((0 << 32) | 1) + O.mix64((2 << 32) | 3)) % (1 << 64)
    assert (h, r, d) == (exp, 2, 3)


# --------------------------------------------------------- brute force ---

def _random_small(rng, n, p):
    edges = [(u, v) for u in range(n) for v in range(n) if rng.random() < p]
    off, adj = _text_to_csr(edges, n)
    w = rng.integers(1, 10, size=len(adj)).astype(np.uint32)
    return O.CSR(off, adj, w), edges


def test_brute_force_tiny_graphs():
    rng = np.random.default_rng(3)
    for trial in range(60):
        n = int(rng.integers(1, 8))
        g, edges = _random_small(rng, n, float(rng.uniform(0.1, 0.6)))
        INF = n + 1
        # Floyd-Warshall on hops and on weights
        hop = np.full((n, n), np.inf)
        wt = np.full((n, n), np.inf)
        for u in range(n):
            hop[u, u] = wt[u, u] = 0
            for i in range(int(g.off[u]), int(g.off[u + 1])):
                v = int(g.adj[i])
                hop[u, v] = min(hop[u, v], 1)
                wt[u, v] = min(wt[u, v], g.w[i])
        for k, i, j in itertools.product(range(n), repeat=3):
            hop[i, j] = min(hop[i, j], hop[i, k] + hop[k, j])
        for k, i, j in itertools.product(range(n), repeat=3):
            wt[i, j] = min(wt[i, j], wt[i, k] + wt[k, j])
        for s in range(n):
            exp_b = [int(min(INF, 1 + h)) if np.isfinite(h) else INF for h in hop[s]]
            exp_s = [int(min(INF, 1 + x)) if np.isfinite(x) else INF for x in wt[s]]
            assert O.bfs(g, s).tolist() == exp_b
            assert O.sssp(g, s).tolist() == exp_s
            assert O.sssp_bf(g, s).tolist() == exp_s
        # CC: least vertex that reaches v (reachability closure)
        exp_c = [min(u for u in range(n) if np.isfinite(hop[u, v])) for v in range(n)]
        assert O.cc(g).tolist() == exp_c
        assert O.cc_minreach(g).tolist() == exp_c


# ---------------------------------------------------------- scipy pins ---

@pytest.fixture(scope="module")
def rmat12w():
    return O.rmat(12, 16, seed=4, weights=True)


def _scipy_matrix(g, weights):
    src = np.repeat(np.arange(g.n), np.diff(g.off).astype(np.int64))
    data = g.w.astype(np.float64) if weights else np.ones(g.m)
    # duplicates: keep the minimum weight (csr_matrix would sum them)
    order = np.lexsort((data, g.adj, src))
    s, d, x = src[order], g.adj[order], data[order]
    keep = np.ones(len(s), bool)
    keep[1:] = (s[1:] != s[:-1]) | (d[1:] != d[:-1])
    return sp.csr_matrix((x[keep], (s[keep], d[keep])), shape=(g.n, g.n))


def _clip(dist, n):
    out = np.full(n, n + 1, np.int64)
    f = np.isfinite(dist)
    out[f] = np.minimum(n + 1, dist[f].astype(np.int64) + 1)
    return out


def test_bfs_sssp_vs_scipy(rmat12w, golden):
    for g in (rmat12w, O.read_graph(os.path.join(golden, "rand64.csr"), "csr")):
        A = _scipy_matrix(g, weights=False)
        Aw = _scipy_matrix(g, weights=True)
        for s in (0, 1, 3, g.n - 1):
            db = csg.shortest_path(A, indices=s, unweighted=True)
            assert np.array_equal(O.bfs(g, s), _clip(db, g.n))
            dw = csg.dijkstra(Aw, indices=s)
            exp = _clip(dw, g.n)
            assert np.array_equal(O.sssp(g, s), exp)
            assert np.array_equal(O.sssp_bf(g, s), exp)
            assert np.array_equal(O.sssp_mt(g, s), exp)
            assert np.array_equal(O.bfs_mt(g, s)[0], O.bfs(g, s))


@pytest.mark.parametrize("delta", [1, 7, 16, 1000])
def test_sssp_delta_mt_vs_scipy(rmat12w, golden, delta):
    """Delta-stepping on the host cores: every bucket width gives scipy's
    Dijkstra distances (clipped at n+1), on the R-MAT graph and the
    converter's random fixture (kept duplicates and self-loops)."""
    for g in (rmat12w, O.read_graph(os.path.join(golden, "rand64.csr"), "csr")):
        Aw = _scipy_matrix(g, weights=True)
        for s in (0, 1, 3, g.n - 1):
            exp = _clip(csg.dijkstra(Aw, indices=s), g.n)
            assert np.array_equal(O.sssp_delta_mt(g, s, delta), exp), (g.n, s, delta)


def test_sssp_delta_mt_clips_at_sentinel():
    """A path 0 -> 1 -> 2 -> 3 with weight 2: 1 + dist passes n + 1 = 5 at
    vertex 2, which then stays at the sentinel (and so does vertex 3 behind
    it), as the strict-less relaxation from val[src] = 1 leaves it."""
    g = O.CSR(np.array([0, 1, 2, 3, 3], np.uint64), np.array([1, 2, 3], np.uint32),
              np.array([2, 2, 2], np.uint32))
    assert O.sssp_delta_mt(g, 0, 16).tolist() == [1, 3, 5, 5]
    assert O.sssp(g, 0).tolist() == [1, 3, 5, 5]


@pytest.mark.parametrize("symmetric", [True, False])
def test_bfs_multi_vs_scipy(symmetric):
    """64 queries at once: each query's value array equals scipy's unweighted
    shortest paths (+1, clipped at n+1), and its digest, reached count and
    depth equal hash_values of that array.  Directed R-MAT runs top-down
    levels only; the symmetric one also bottom-up levels."""
    g = O.rmat(12, 16, seed=9, symmetric=symmetric)
    A = _scipy_matrix(g, weights=False)
    rng = np.random.default_rng(2)
    srcs = rng.integers(0, g.n, 64).astype(np.uint32)
    srcs[1] = srcs[2]                                     # a repeated source
    srcs[3] = int(np.flatnonzero(np.diff(g.off) == 0)[0])  # an isolated one
    out = O.bfs_multi(g, srcs, symmetric=symmetric, full=True)
    for q, s in enumerate(srcs):
        exp = _clip(csg.shortest_path(A, indices=int(s), unweighted=True), g.n)
        assert np.array_equal(out["full"][q], exp), (q, s)
        h, r, d = O.hash_values(exp.astype(np.uint32), g.n + 1)
        assert (out["hash"][q], out["reached"][q], out["depth"][q]) == (h, r, d), (q, s)


def test_bfs_multi_small_k_and_ragged_n():
    """k below 64 and vertex counts that are not multiples of 64: equal to
    the serial queue BFS query by query."""
    rng = np.random.default_rng(5)
    for n in (1, 5, 65, 200):
        g, _ = _random_small(rng, n, 0.05) if n > 1 else _random_small(rng, 1, 0.0)
        for k in (1, 3, 17):
            srcs = rng.integers(0, n, k).astype(np.uint32)
            out = O.bfs_multi(g, srcs, full=True)
            for q, s in enumerate(srcs):
                assert np.array_equal(out["full"][q], O.bfs(g, int(s))), (n, k, q)


def test_cc_vs_scipy_symmetric(rmat12w):
    g = rmat12w
    ncomp, lab = csg.connected_components(_scipy_matrix(g, False), directed=False)
    first = np.full(ncomp, g.n, np.int64)
    np.minimum.at(first, lab, np.arange(g.n))
    exp = first[lab]
    assert np.array_equal(O.cc(g), exp)
    assert np.array_equal(O.cc_mt(g), exp)
    assert np.array_equal(O.cc_minreach(g), exp)


def test_cc_directed_algorithms_agree():
    g = O.rmat(11, 8, seed=6, symmetric=False)
    a = O.cc(g)
    assert np.array_equal(a, O.cc_minreach(g))
    assert np.array_equal(a, O.cc_mt(g))


def test_pagerank_vs_sparse_solve():
    g = O.rmat(10, 8, seed=8, symmetric=False)
    alpha = O.ALPHA_F32
    deg = np.diff(g.off).astype(np.float64)
    src = np.repeat(np.arange(g.n), deg.astype(np.int64))
    # M[v, u] = 1/outdeg(u) for every edge u -> v (duplicates add up)
    M = sp.csr_matrix((1.0 / deg[src], (g.adj, src)), shape=(g.n, g.n))
    r = spla.spsolve((sp.identity(g.n) - alpha * M).tocsc(), np.full(g.n, 1 - alpha))
    ex, _ = O.pr_exact(g)
    assert np.max(np.abs(ex - r)) < 1e-9
    exm, _ = O.pr_exact_mt(O.transpose(g), tol=1e-13)
    assert np.max(np.abs(exm - r)) < 1e-9
    # the reference's gated loops converge to the same fixpoint
    push, _ = O.pr_push_ref(g, tol=1e-13, max_iter=100000)
    pull, _ = O.pr_pull_ref(O.transpose(g), tol=1e-13, max_iter=100000)
    assert np.max(np.abs(push - r)) < 1e-8
    assert np.max(np.abs(pull - r)) < 1e-8
    # fp32 CPU baseline within the GPU bar
    p32, _ = O.pr_push_mt(g)
    q32, _ = O.pr_pull_mt(O.transpose(g))
    assert O.normalized_l1(p32, r) < 1e-5
    assert O.normalized_l1(q32, r) < 1e-5


# ------------------------------------------------------------- R-MAT ---

def test_rmat_structure():
    g = O.rmat(12, 16, seed=1, weights=True)
    n = g.n
    src = np.repeat(np.arange(n), np.diff(g.off).astype(np.int64))
    assert not np.any(src == g.adj)                          # no self-loops
    key = src.astype(np.uint64) << np.uint64(32) | g.adj.astype(np.uint64)
    assert np.all(np.diff(key) > 0)                          # sorted, unique
    rev = g.adj.astype(np.uint64) << np.uint64(32) | src.astype(np.uint64)
    assert np.array_equal(np.sort(rev), key)                 # symmetric
    assert g.w.min() >= 1 and g.w.max() <= 64
    # weight symmetric: w(u,v) == w(v,u)
    idx = np.searchsorted(key, rev)
    assert np.array_equal(g.w, g.w[idx])
    # numpy and C weight functions agree
    for i in (0, 17, g.m - 1):
        assert O.lib().orc_weight(1, int(src[i]), int(g.adj[i])) == g.w[i]
    # degree skew of R-MAT: vertex 0 is the heaviest
    deg = np.diff(g.off)
    assert deg[0] == deg.max()


def test_rmat_quadrant_frequencies():
    th = np.zeros(3, np.uint64)
    O.lib().orc_rmat_thresholds(0.57, 0.19, 0.19, th)
    N = 200000
    u = np.zeros(N, np.uint32)
    v = np.zeros(N, np.uint32)
    O.lib().orc_rmat_tuples(1, 1, th, 0, N, u, v)  # scale 1: one quadrant draw
    q = 2 * u + v
    freq = np.bincount(q, minlength=4) / N
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)
