"""Plain device copy of a known size, the yardstick for HBM byte counters.

    python water-simulator_amd/copy_calibration.py [MiB]

Copies an fp64 buffer (default 1024 MiB) with torch's elementwise copy kernel,
five times, so a counter run (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) can be
scaled: expected bytes per dispatch = MiB * 2**20 read and the same written.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x5", flush=True)


if __name__ == "__main__":
    main()
