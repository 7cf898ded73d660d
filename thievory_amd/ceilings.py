"""Link and memory ceilings of the device the engine runs on.

    python -m thievory_amd.ceilings [--bytes N]

Prints one JSON line: the rates thv_link_probe measures (zero-copy kernel
read of mapped pinned host memory, pinned host-to-device DMA, HBM read by
the same streaming kernel, HBM copy counting bytes read and written) and
the device.  Under `rocprofv3 --pmc` its `k_stream_read` dispatch over the
host buffer reads a known byte count, which calibrates the IO read-request
counter (TCC_EA0_RDREQ_IO_32B) the host-link measurements use.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import thievory_amd as T  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 30)
    a = ap.parse_args(argv)
    r = T.link_probe(a.bytes)
    r["device"] = T.device_info()
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
