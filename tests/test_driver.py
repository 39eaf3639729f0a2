"""The drop-in driver: girih's mwd_kernel command line on the GPU target.

CPU tests cover argument handling and girih's validity errors (no device call is
made before them); GPU tests run --verify (girih's exact serial comparison) and a
performance run whose report carries the keys girih's scripts/parse.py scrapes.
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = {p: os.path.join(REPO, "girih_amd", "driver", "bin", "mwd_kernel_gpu_" + p) for p in ("dp", "sp")}


def run(prec, *args, timeout=120):
    exe = BIN[prec]
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def test_list_and_help():
    r = run("dp", "--list")
    assert r.returncode == 0
    assert "02   Diamond" in r.stdout and "03   MI355X" in r.stdout
    assert "05  stencil_op:star  time-order:1  radius:1  coeff:variable no-symmetry" in r.stdout
    assert "Relaxed synchronization wavefront" in r.stdout
    r = run("sp", "--help")
    assert r.returncode == 0 and "--target-ts" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--npx", "2"], "topology"),
    (["--target-kernel", "6"], "kernel not available"),
    (["--target-ts", "7"], "unknown time stepper"),
    (["--target-ts", "2", "--target-kernel", "0", "--mwd-type", "2"], "Relaxed synchronization"),
    (["--nx", "0"], "empty domain"),
])
def test_girih_validity_errors(args, msg):
    r = run("dp", *args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_invalid_argument_prints_usage():
    r = run("dp", "--bogus-flag", "3")
    assert "Invalid arguments" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("case", [
    ["--target-ts", 0, "--target-kernel", 0, "--nx", 16, "--ny", 32, "--nz", 24, "--nt", 10],
    ["--target-ts", 2, "--target-kernel", 0, "--nx", 16, "--ny", 32, "--nz", 40, "--nt", 9, "--t-dim", 1,
     "--thread-group-size", 1],
    ["--target-ts", 2, "--target-kernel", 1, "--nx", 70, "--ny", 32, "--nz", 24, "--nt", 9, "--t-dim", 3],
    ["--target-ts", 3, "--target-kernel", 5, "--nx", 33, "--ny", 17, "--nz", 12, "--nt", 7],
    ["--target-ts", 3, "--target-kernel", 1, "--nx", 70, "--ny", 40, "--nz", 30, "--nt", 11],
    ["--target-ts", 3, "--target-kernel", 0, "--nx", 72, "--ny", 40, "--nz", 36, "--nt", 8],
    ["--target-ts", 1, "--target-kernel", 4, "--nx", 20, "--ny", 12, "--nz", 10, "--nt", 4],
    ["--target-ts", 0, "--target-kernel", 2, "--nx", 20, "--ny", 12, "--nz", 10, "--nt", 5],
    ["--target-ts", 0, "--target-kernel", 3, "--nx", 20, "--ny", 12, "--nz", 10, "--nt", 5],
])
def test_verify_passes(prec, case):
    print("driver", prec, case, flush=True)
    r = run(prec, "--verify", 1, "--verbose", 0, *case)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"eMax:0\.000e\+00\|eL1:0\.000e\+00-PASSED", r.stdout), r.stdout


@pytest.mark.gpu
def test_diamond_rejects_non_multiple_width():
    r = run("dp", "--verify", 1, "--target-ts", 2, "--target-kernel", 0, "--nx", 16, "--ny", 24, "--nz", 40,
            "--nt", 9, "--t-dim", 1)
    assert r.returncode == 1 and "diamond width" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("ts", [0, 2, 3])
def test_performance_report_keys(ts):
    r = run("sp", "--target-ts", ts, "--target-kernel", 0, "--nx", 64, "--ny", 64, "--nz", 64, "--nt", 20,
            "--t-dim", 1, "--n-tests", 3)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    for key in ("Time stepper name:", "Stencil Kernel name: star", "Stencil Kernel semi-bandwidth: 4",
                "Precision: SP", "Number of time steps: 20", "Number of tests: 3", "COMPLETED SUCCESSFULLY"):
        assert key in out, key
    if ts == 2:
        assert re.search(r"MWD main-loop RANK0 MStencil/s MAX: [0-9.]+", out)
        assert "Intra-diamond width:   16" in out
    else:
        assert re.search(r"RANK0 GStencil/s MEDIAN: [0-9.]+", out)
