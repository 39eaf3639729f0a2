"""Known-size streaming kernels used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE.

    python -m girih_amd.calib [--mib 1024] [--reps 5] [--time 50]

Runs `dst = src * 1` (torch's elementwise kernel: one plain vector load and one
store per element) on a --mib MiB buffer --reps times.  Each dispatch reads and
writes exactly that many bytes, so the counter reading per dispatch divided by
the buffer size is the factor that turns the counters into bytes.

With --time K it also times K more copies (HIP events) and prints the copy rate,
(bytes read + bytes written) / time: the HBM throughput a plain streaming
kernel reaches, next to the published 8 TB/s.  --triad also times `a = b + c`
(torch's elementwise add), two streams read per stream written: the read/write
mix of a wave sweep (u^n and u^(n-1) read, u^(n+1) written), whose rate the
sweep's algorithmic 12 B per lattice update is compared with.
"""
import json
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--time", type=int, default=0, help="timed copies after the calibration dispatches")
    ap.add_argument("--triad", action="store_true", help="also time a = b + c on --mib MiB arrays")
    a = ap.parse_args()
    n = a.mib * 1024 * 1024 // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("streamed %d MiB x %d" % (a.mib, a.reps))
    if a.time > 0:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.time):
            torch.mul(src, 1.0, out=dst)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / a.time
        print(json.dumps({"copy_mib": a.mib, "copies": a.time, "ms_per_copy": round(ms, 4),
                          "copy_gbs": round(2 * a.mib * 1048576 / ms / 1e6, 1)}))
        if a.triad:
            c = torch.rand(n, device="cuda")
            torch.add(src, c, out=dst)
            e0.record()
            for _ in range(a.time):
                torch.add(src, c, out=dst)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / a.time
            print(json.dumps({"triad_mib": a.mib, "triads": a.time, "ms_per_triad": round(ms, 4),
                              "triad_gbs": round(3 * a.mib * 1048576 / ms / 1e6, 1)}))


if __name__ == "__main__":
    main()
