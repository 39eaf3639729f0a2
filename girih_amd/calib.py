"""Known-size copy used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE readings.

    python -m girih_amd.calib [--mib 1024] [--reps 5]

Copies a --mib MiB buffer into another one --reps times (torch's device copy
kernel).  Each dispatch moves exactly that many bytes in and out of HBM, so the
counter reading per dispatch divided by the buffer size is the scale factor that
turns the counters into bytes.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    n = a.mib * 1024 * 1024 // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copied %d MiB x %d" % (a.mib, a.reps))


if __name__ == "__main__":
    main()
