"""GPU checks of the 64-bit-id path (thv_graph64_*, thv_bfs64 / cc64 /
sssp64): a seeded R-MAT graph is spread over a larger id space by the
order-preserving map id(i) = base + i * STRIDE, so BFS levels, SSSP
distances and CC least-id labels of the mapped vertices are the oracle's
on the compact graph (mapped), and every other vertex is unreached (CC: its
own label).  The largest case has about 2^32 + 2^27 vertices, with mapped ids
on both sides of 2^32; its values are checked through the device digest,
the reached count, the depth and a gather of the mapped vertices."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _say(*a):
    print(*a, flush=True)


def mix64(z):
    """SplitMix64's output function (the digest's mixer), on uint64 arrays."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def digest(ids, vals):
    with np.errstate(over="ignore"):
        return int(np.sum(mix64(mix64(ids) + vals), dtype=np.uint64))


def spread(og, n, base, stride, weights):
    """The 64-bit lists of og with vertex i at base + i * stride."""
    deg = np.diff(og.off).astype(np.int64)
    listed = np.flatnonzero(deg > 0)
    ids = base + listed.astype(np.uint64) * np.uint64(stride)
    loff = np.zeros(len(listed) + 1, np.uint64)
    loff[1:] = np.cumsum(deg[listed])
    dst = base + og.adj.astype(np.uint64) * np.uint64(stride)
    w = og.w if weights else None
    return T.Graph64(n, ids, loff, dst, w)


def mapped_values(ref, n_small, n_big, base, stride, ident=False):
    """Compact values to the big graph's: levels / distances stay, the
    compact sentinel becomes n + 1; labels (ident) are ids and map."""
    ref = ref.astype(np.uint64)
    if ident:
        return base + ref * np.uint64(stride)
    return np.where(ref == n_small + 1, np.uint64(n_big + 1), ref)


@pytest.mark.parametrize("budget_frac", [1.0, 0.25])
def test_wide_dense_ids_full_output(budget_frac):
    """R-MAT 12 / ef 8 with weights at ids 3 + 17 i in a 2^17-vertex graph:
    the whole value arrays equal the mapped oracle results."""
    _say("case: wide, R-MAT 12 in 2^17 ids, budget", budget_frac)
    og = O.rmat(12, 8, seed=6, weights=True)
    n, base, stride = 1 << 17, 3, 17
    g = spread(og, n, base, stride, True)
    g.store_config(int(budget_frac * g.m * 16), 256)
    info = g.store_info()
    assert (info["device_edges"] == g.m) if budget_frac == 1.0 else (info["device_edges"] < g.m)
    pos = base + np.arange(og.n, dtype=np.uint64) * np.uint64(stride)
    for algo in ("bfs", "sssp", "cc"):
        r = g.cc(out=True) if algo == "cc" else getattr(g, algo)(int(pos[1]), out=True)
        if algo == "cc":
            exp = np.arange(n, dtype=np.uint64)
            exp[pos] = mapped_values(O.cc(og), og.n, n, base, stride, ident=True)
        else:
            ref = O.bfs(og, 1) if algo == "bfs" else O.sssp(og, 1)
            exp = np.full(n, n + 1, np.uint64)
            exp[pos] = mapped_values(ref, og.n, n, base, stride)
        assert np.array_equal(r["values"], exp), algo
        fin = exp != n + 1 if algo != "cc" else np.ones(n, bool)
        ids = np.flatnonzero(fin).astype(np.uint64)
        assert r["reached"] == len(ids)
        assert r["hash"] == digest(ids, exp[fin])
        assert r["depth"] == int(exp[fin].max())


def test_wide_past_2_32():
    """The mapped R-MAT 12 vertices sit at ids 2^32 - 2048 * 65537 + 65537 i,
    on both sides of 2^32; n is 12345 past the last (about 2^32 + 2^27)."""
    og = O.rmat(12, 8, seed=9, weights=True)
    stride = 65537
    base = (1 << 32) - 2048 * stride
    pos = base + np.arange(og.n, dtype=np.uint64) * np.uint64(stride)
    n = int(pos[-1]) + 12345
    assert pos[0] < (1 << 32) < pos[-1] < n
    _say("case: wide, n =", n)
    g = spread(og, n, base, stride, True)
    deg = np.diff(og.off)
    src = int(np.flatnonzero(deg > 0)[-1])       # a source above 2^32
    assert pos[src] > (1 << 32)
    for algo in ("bfs", "sssp"):
        ref = O.bfs(og, src) if algo == "bfs" else O.sssp(og, src)
        exp = mapped_values(ref, og.n, n, base, stride)
        r = getattr(g, algo)(int(pos[src]))
        fin = exp != n + 1
        assert r["reached"] == int(fin.sum()), algo
        assert r["hash"] == digest(pos[fin], exp[fin]), algo
        assert r["depth"] == int(exp[fin].max()), algo
        assert np.array_equal(g.values(pos), exp), algo
        # a vertex between two mapped ones and the last vertex: unreached
        assert g.values([pos[src] + 1, n - 1]).tolist() == [n + 1, n + 1]
    r = g.cc()
    assert r["reached"] == n                     # every vertex has a label
    lab = mapped_values(O.cc(og), og.n, n, base, stride, ident=True)
    assert np.array_equal(g.values(pos), lab)
    assert g.values([5, pos[3] + 1, n - 1]).tolist() == [5, int(pos[3]) + 1, n - 1]


def test_wide_refusals():
    _say("case: wide refusals")
    with pytest.raises(T.ThvError):   # ids not ascending
        T.Graph64(10, [3, 2], [0, 1, 2], [2, 3])
    with pytest.raises(T.ThvError):   # destination out of range
        T.Graph64(10, [2, 3], [0, 1, 2], [3, 10])
    with pytest.raises(T.ThvError):   # empty list
        T.Graph64(10, [2, 3], [0, 0, 1], [3])
    g = T.Graph64(10, [2, 3], [0, 1, 2], [3, 2])
    with pytest.raises(T.ThvError) as e:   # no weights
        g.sssp(2)
    assert e.value.code == -7
    with pytest.raises(T.ThvError):
        g.bfs(10)
    r = g.bfs(2, out=True)
    assert r["values"].tolist() == [11, 11, 1, 2, 11, 11, 11, 11, 11, 11]
