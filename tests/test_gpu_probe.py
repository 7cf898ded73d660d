"""GPU check of the ceilings probe (thv_link_probe): every rate is positive,
the zero-copy read of host memory is slower than the HBM read, and a buffer
under 1 MiB is refused."""
import pytest

import thievory_amd as T

pytestmark = pytest.mark.gpu


def test_link_probe():
    print("case: link probe 64 MiB", flush=True)
    r = T.link_probe(64 << 20)
    for k in ("zero_copy_read_GBps", "dma_h2d_GBps", "hbm_read_GBps", "hbm_copy_GBps"):
        assert r[k] > 0, k
    assert r["zero_copy_read_GBps"] < r["hbm_read_GBps"]
    with pytest.raises(T.ThvError) as e:
        T.link_probe(1 << 16)
    assert e.value.code == -1
