"""GPU checks of thv_bfs_noalloc, the BFS whose per-call buffers live in the
library's static device array: its levels equal thv_bfs's (and the
oracle's) bit for bit on two calls in a row, the second on a smaller graph,
then once more on the first graph (the array keeps the last call's
contents); it refuses a graph whose workspace exceeds the array and a graph
whose store is not built."""
import numpy as np
import pytest

from oracle import oracle as O
import thievory_amd as T

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20  # THV_NOALLOC_SCRATCH_BYTES


def _say(*a):
    print(*a, flush=True)


def test_noalloc_matches_bfs_two_calls():
    _say("case: noalloc BFS, R-MAT 18 split store then R-MAT 13 resident")
    big = T.Graph.rmat(18, 16, seed=5)
    big.store_config(int(0.25 * big.m * 4), 1 << 12)
    small = T.Graph.rmat(13, 8, seed=9, symmetric=False)
    small.store_config(None, 1 << 16)
    og_big, og_small = O.rmat(18, 16, seed=5), O.rmat(13, 8, seed=9, symmetric=False)
    for g, og, src in ((big, og_big, 1), (small, og_small, 3), (big, og_big, 7)):
        ref, _ = g.bfs(src)
        lv, st = g.bfs_noalloc(src)
        assert np.array_equal(lv, ref), (g.n, src)
        assert np.array_equal(lv, O.bfs(og, src)), (g.n, src)
        assert st.reached == int(np.count_nonzero(ref != g.n + 1))
        assert st.iterations == int(ref[ref != g.n + 1].max())


def test_noalloc_refuses_what_it_cannot_hold():
    _say("case: noalloc BFS limits")
    g = T.Graph.rmat(12, 8, seed=2)
    with pytest.raises(T.ThvError) as e:  # store not built yet
        g.bfs_noalloc(1)
    assert e.value.code == -8
    big = T.Graph.rmat(22, 4, seed=2)  # ~30 bytes per vertex: over 64 MiB
    big.store_config(None, 1 << 16)
    with pytest.raises(T.ThvError) as e:
        big.bfs_noalloc(1)
    assert e.value.code == -1
    # the refusal left nothing behind: a graph that fits still runs
    g.store_config(None, 1 << 16)
    lv, _ = g.bfs_noalloc(1)
    assert np.array_equal(lv, g.bfs(1)[0])
