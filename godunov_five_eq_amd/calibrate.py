"""A plain device copy of known size, for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters
into bytes: each of the timed launches reads and writes exactly `bytes` bytes.

    rocprofv3 --pmc FETCH_SIZE -f csv -d DIR -o copy -- python -m godunov_five_eq_amd.calibrate
"""
import json

import torch


def main(nbytes=1 << 30, launches=3):
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": launches}))


if __name__ == "__main__":
    main()
