"""Plain copy of a known size, for scaling memory-traffic counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 thievory_amd/calibrate.py
    rocprofv3 --pmc WRITE_SIZE -- python3 thievory_amd/calibrate.py

Copies 1 GiB device-to-device three times (torch's elementwise copy
kernel).  An ideal counter reads 1,048,576 KiB fetched and written per
dispatch; the ratio of the measured value to that is the factor applied to
the traversal kernels' counters.
"""
import torch

N = 1 << 28  # int32 elements = 1 GiB


def main():
    a = torch.ones(N, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {a.numel() * 4} bytes x3", flush=True)


if __name__ == "__main__":
    main()
