"""Reference copy for HBM traffic counters: a device-to-device copy of a known size.

Run under ``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``): the counter per copy dispatch over
the bytes copied is the factor that turns the counter into bytes for the kernels measured in the
same way. Prints the size of each copy.
"""
from __future__ import annotations

import sys


def main(mib: int = 1024, reps: int = 3) -> int:
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {reps} x {mib} MiB ({n * 4} bytes each)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(*(int(a) for a in sys.argv[1:])))
