"""Known-size streaming kernels used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE.

    python -m girih_amd.calib [--mib 1024] [--reps 5]

Runs `dst = src * 1` (torch's elementwise kernel: one plain vector load and one
store per element) on a --mib MiB buffer --reps times.  Each dispatch reads and
writes exactly that many bytes, so the counter reading per dispatch divided by
the buffer size is the factor that turns the counters into bytes.
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
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("streamed %d MiB x %d" % (a.mib, a.reps))


if __name__ == "__main__":
    main()
