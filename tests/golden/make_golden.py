"""Capture golden vectors of girih's own steppers for the CPU oracle and the HIP path.

Runs oracle/_ref/ref_dump_{dp,sp} (girih's C sources compiled where they lie by
`make -C oracle ref`, linked with oracle/ref_dump.c) over a reduced copy of the
reference's verification matrix (scripts/verification/verification_std.py:3-96,
verification_idiam.py:3-154): spatial blocking (ts 0) for kernels 0-5 and the
diamond stepper (ts 2; 1WD and multi-thread-group MWD types 0-3) for kernels 0, 1, 5.
Every run also passes girih's own exact --verify comparison (ref_dump checks it).

Writes tests/golden/ref_cases.json: per case the flags, the applied nt (after the
diamond stepper's rounding), the padded domain shape, SHA-256 of U1 and U2 in the
reference's padded layout and a few summary values; and tests/golden/ref_arrays.npz
with the full U1/U2 of the smallest cases.

usage: python tests/golden/make_golden.py   (needs /root/reference and `make -C oracle ref`)
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")

# (name, precision, nx, ny, nz, nt, ts, kernel, extra flags, OMP threads)
SB = ["--thread-group-size", "2"]


def wd(t_dim, tgs=1, thz=1, nwf=1, mwd_type=0):
    return ["--thread-group-size", str(tgs), "--thx", "1", "--thy", "1", "--thz", str(thz),
            "--t-dim", str(t_dim), "--num-wavefronts", str(nwf), "--mwd-type", str(mwd_type)]


CASES = []
for prec in ("dp", "sp"):
    CASES += [
        ("sb_k0_16x32x24_nt10", prec, 16, 32, 24, 10, 0, 0, SB, 2),
        ("sb_k0_24x16x20_nt9", prec, 24, 16, 20, 9, 0, 0, SB, 2),
        ("sb_k1_16x32x24_nt10", prec, 16, 32, 24, 10, 0, 1, SB, 2),
        ("sb_k1_20x12x16_nt7", prec, 20, 12, 16, 7, 0, 1, SB, 2),
        ("sb_k2_16x32x24_nt10", prec, 16, 32, 24, 10, 0, 2, SB, 2),
        ("sb_k3_16x32x24_nt10", prec, 16, 32, 24, 10, 0, 3, SB, 2),
        ("sb_k4_16x32x24_nt6", prec, 16, 32, 24, 6, 0, 4, SB, 2),
        ("sb_k5_16x32x24_nt10", prec, 16, 32, 24, 10, 0, 5, SB, 2),
        ("sb_k5_13x9x11_nt5", prec, 13, 9, 11, 5, 0, 5, SB, 2),
        ("sb_k0_8x8x8_nt4", prec, 8, 8, 8, 4, 0, 0, SB, 2),
        ("1wd_k0_td1_16x32x40_nt9", prec, 16, 32, 40, 9, 2, 0, wd(1), 2),
        ("1wd_k0_td3_16x64x40_nt12", prec, 16, 64, 40, 12, 2, 0, wd(3), 2),
        ("1wd_k1_td1_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 1, wd(1), 2),
        ("1wd_k1_td3_16x32x24_nt9", prec, 16, 32, 24, 9, 2, 1, wd(3), 2),
        ("1wd_k5_td3_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 5, wd(3), 2),
        ("mwd0_k0_td1_tgs2_16x32x40_nt10", prec, 16, 32, 40, 10, 2, 0, wd(1, 2, 2, 2, 0), 4),
        ("mwd1_k0_td1_tgs2_16x32x40_nt10", prec, 16, 32, 40, 10, 2, 0, wd(1, 2, 2, 2, 1), 4),
        ("mwd0_k1_td3_tgs2_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 1, wd(3, 2, 2, 2, 0), 4),
        ("mwd2_k1_td3_tgs2_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 1, wd(3, 2, 2, 2, 2), 4),
        ("mwd3_k1_td3_tgs2_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 1, wd(3, 2, 2, 2, 3), 4),
        ("mwd1_k5_td1_tgs2_16x32x24_nt10", prec, 16, 32, 24, 10, 2, 5, wd(1, 2, 2, 2, 1), 4),
    ]

# full arrays are stored for these (small) cases
FULL = {"sb_k0_8x8x8_nt4", "sb_k1_20x12x16_nt7", "sb_k5_13x9x11_nt5"}


def read_dump(path):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(48), dtype=np.int32)
        assert hdr[0] == 0x47524846, "bad magic"
        dt = np.float64 if hdr[1] == 8 else np.float32
        nnx, nny, nnz = int(hdr[6]), int(hdr[7]), int(hdr[8])
        n = nnx * nny * nnz
        u1 = np.frombuffer(f.read(n * hdr[1]), dtype=dt).reshape(nnz, nny, nnx)
        u2 = np.frombuffer(f.read(n * hdr[1]), dtype=dt).reshape(nnz, nny, nnx)
    return hdr, u1, u2


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest()


def main():
    out_cases = []
    arrays = {}
    with tempfile.TemporaryDirectory() as td:
        for (name, prec, nx, ny, nz, nt, ts, k, extra, nthr) in CASES:
            exe = os.path.join(REF_DIR, "ref_dump_" + prec)
            dump = os.path.join(td, "d.bin")
            flags = ["--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--nt", str(nt),
                     "--target-ts", str(ts), "--target-kernel", str(k), "--verbose", "0"] + extra
            env = dict(os.environ, OMP_NUM_THREADS=str(nthr))
            r = subprocess.run([exe, dump] + flags, env=env, capture_output=True, text=True, timeout=300)
            if r.returncode != 0 or "PASSED" not in r.stdout:
                sys.exit("reference run failed for %s:\n%s\n%s" % (name, r.stdout, r.stderr))
            hdr, u1, u2 = read_dump(dump)
            case = {
                "name": prec + "_" + name, "precision": prec, "nx": nx, "ny": ny, "nz": nz,
                "nt_flag": nt, "nt_applied": int(hdr[9]), "ts": ts, "kernel": k, "radius": int(hdr[5]),
                "ldomain_shape": [int(hdr[6]), int(hdr[7]), int(hdr[8])],
                "flags": flags, "omp_threads": nthr,
                "sha256_u1": sha(u1), "sha256_u2": sha(u2),
                "sum_u1": float(np.sum(u1, dtype=np.float64)), "sum_u2": float(np.sum(u2, dtype=np.float64)),
                "ref_verify": r.stdout.strip().splitlines()[-1],
            }
            out_cases.append(case)
            if name in FULL:
                arrays[case["name"] + "__u1"] = u1.copy()
                arrays[case["name"] + "__u2"] = u2.copy()
            print("%-44s nt %3d -> %3d  %s" % (case["name"], nt, case["nt_applied"], case["ref_verify"]))
    with open(os.path.join(HERE, "ref_cases.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "source": "girih src/ compiled by oracle/Makefile (ref target), captured by oracle/ref_dump.c",
                   "cases": out_cases}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_arrays.npz"), **arrays)


if __name__ == "__main__":
    main()
