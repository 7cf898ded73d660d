"""GPU check of the per-partition activity filter (common.cuh:120-183 skips
host partitions with no active vertex): with every partition in host memory,
each top-down level reads exactly the partitions that hold a frontier
vertex with edges, and skips the rest.  Partitions are the greedy vertex
ranges of at most E edges (one oversize vertex alone, graph.cuh:270-304);
the expected counts come from the oracle's levels."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu


def greedy_partitions(off, E):
    """First vertex of every partition, and n at the end."""
    deg = np.diff(off).astype(np.int64)
    pv, cur = [0], 0
    for v, d in enumerate(deg):
        if cur + d <= E or v == pv[-1]:
            cur += d
        else:
            pv.append(v)
            cur = d
    pv.append(len(deg))
    return np.array(pv, np.int64)


@pytest.mark.parametrize("symmetric", [False, True])
def test_activity_filter_counts(symmetric):
    print("case: activity filter, symmetric", symmetric, flush=True)
    og = O.rmat(14, 8, seed=5, symmetric=symmetric)
    g = T.Graph.rmat(14, 8, seed=5, symmetric=symmetric)
    g.set_option("direction", 0)          # top-down levels only
    E = 1 << 10
    g.store_config(0, E)                  # every partition on the host
    info = g.store_info()
    assert info["device_partitions"] == 0 and info["device_edges"] == 0
    pv = greedy_partitions(og.off, E)
    P = len(pv) - 1
    assert info["partitions"] == P
    deg = np.diff(og.off).astype(np.int64)
    part = np.searchsorted(pv, np.arange(og.n), side="right") - 1
    for s in (1, 77, 9000):
        ref = O.bfs(og, s)
        lv, st = g.bfs(s)
        assert np.array_equal(lv, ref), s
        touched, levels = 0, 0
        while True:
            front = (ref == levels + 1) & (deg > 0)
            if not front.any():
                break
            touched += len(np.unique(part[front]))
            levels += 1
        assert st.iterations == levels, s
        assert st.host_partitions_touched == touched, s
        assert st.host_partitions_skipped == levels * P - touched, s
        assert 0 < touched < levels * P, s    # the filter skipped partitions
