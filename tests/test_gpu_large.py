"""GPU test of a graph with more than 2^32 edge records.

The Hamming graph H(5, 32): n = 32^5 = 2^25 vertices, each adjacent to the
155 vertices that differ from it in exactly one base-32 digit, so
m = 155 * 2^25 = 5,200,936,960 > 2^32.  Its BFS levels have a closed form:
1 + the number of base-32 digits in which v differs from the source.  The
graph is connected, so every CC label is 0.  This exercises 64-bit edge
offsets, queue edge prefixes, chunk tables and counters with a quarter of
the records in HBM and the rest read zero-copy."""
import numpy as np
import pytest

import thievory_amd as T

pytestmark = pytest.mark.gpu

DIGITS, BITS = 5, 5


def _say(*a):
    print(*a, flush=True)


def _hamming_csr():
    n = 1 << (DIGITS * BITS)
    masks = np.array([t << (BITS * p) for p in range(DIGITS) for t in range(1, 1 << BITS)],
                     np.uint32)
    deg = len(masks)
    dst = np.empty(n * deg, np.uint32)
    step = 1 << 20
    for a in range(0, n, step):
        v = np.arange(a, a + step, dtype=np.uint32)
        dst[a * deg:(a + step) * deg] = (v[:, None] ^ masks[None, :]).ravel()
    off = np.arange(n + 1, dtype=np.uint64) * np.uint64(deg)
    return n, off, dst


def _levels(n, src):
    x = np.arange(n, dtype=np.uint32) ^ np.uint32(src)
    diff = np.zeros(n, np.uint32)
    for p in range(DIGITS):
        diff += ((x >> np.uint32(BITS * p)) & np.uint32((1 << BITS) - 1)) != 0
    return diff + 1


@pytest.mark.timeout(900)
def test_more_than_2_32_edges():
    _say("case: hamming graph H(5, 32), m > 2^32")
    n, off, dst = _hamming_csr()
    m = len(dst)
    assert m > 1 << 32
    g = T.Graph.from_arrays(off, dst)
    del dst
    assert (g.n, g.m) == (n, m)
    g.set_option("symmetric", 1)
    g.store_config(int(0.25 * m * 4), 1 << 16)
    _say("store configured")
    src = 0x1234567 & (n - 1)
    want = _levels(n, src)
    lev, st = g.bfs(src)
    assert np.array_equal(lev, want)
    g.set_option("direction", 0)
    lev_td, st_td = g.bfs(src)
    g.set_option("direction", 1)
    assert np.array_equal(lev_td, want)
    # a top-down traversal scans every list of every reached vertex: all m
    assert st_td.edges_traversed == m
    _say("bfs ok", st.ms, st_td.ms)
    lab, _ = g.cc()
    assert not lab.any()
    g.close()
