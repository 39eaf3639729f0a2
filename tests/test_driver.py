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
    nt = 22 if ts == 2 else 20  # the diamond stepper rounds nt at initialisation (diamond_utils.c:1042-1056)
    for key in ("Time stepper name:", "Stencil Kernel name: star", "Stencil Kernel semi-bandwidth: 4",
                "Precision: SP", "Number of time steps: %d" % nt, "Number of tests: 3", "COMPLETED SUCCESSFULLY"):
        assert key in out, key
    if ts == 2:
        assert "###INFO: Modified nt from 020 to 022" in out
    if ts == 2:
        assert re.search(r"MWD main-loop RANK0 MStencil/s MAX: [0-9.]+", out)
        assert "Intra-diamond width:   16" in out
    else:
        assert re.search(r"RANK0 GStencil/s MEDIAN: [0-9.]+", out)


REF_BIN = os.path.join(REPO, "oracle", "_ref", "mwd_kernel_sp")


def _report_keys(text):
    """The 'Key' of every 'Key: value' line of a girih report (what scripts/parse.py
    matches on), with the time stamps of the first and last lines dropped."""
    keys = set()
    for line in text.splitlines():
        if ":" not in line or line.startswith(("libgomp", "[AUTO TUNE]")):
            continue
        key = line.split(":", 1)[0].rstrip()
        for stamp in ("Started on", "COMPLETED SUCCESSFULLY on"):
            if key.startswith(stamp):
                key = stamp
        keys.add(key)
    return keys


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    ["--target-ts", 0, "--target-kernel", 1, "--nx", 64, "--ny", 32, "--nz", 32, "--nt", 10,
     "--thread-group-size", 1],
    ["--target-ts", 2, "--target-kernel", 0, "--nx", 32, "--ny", 64, "--nz", 32, "--nt", 10, "--t-dim", 1,
     "--thread-group-size", 1, "--thx", 1, "--thy", 1, "--thz", 1, "--num-wavefronts", 1],
    ["--target-ts", 2, "--target-kernel", 5, "--nx", 32, "--ny", 64, "--nz", 32, "--nt", 10, "--t-dim", 3,
     "--thread-group-size", 2, "--thx", 1, "--thy", 1, "--thz", 2, "--num-wavefronts", 2, "--mwd-type", 1],
])
def test_report_has_every_key_of_girihs_report(case):
    """Every 'Key:' line girih's own mwd_kernel prints for these flags (built from its
    sources into oracle/_ref) also appears in the GPU driver's report, so that
    scripts/parse.py reads GPU runs like CPU runs (utils.c:866-1065)."""
    if not os.path.exists(REF_BIN):
        pytest.skip("girih's mwd_kernel is not built (oracle/Makefile ref)")
    args = [str(a) for a in case] + ["--n-tests", "3"]
    ref = subprocess.run([REF_BIN] + args, capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert ref.returncode == 0, ref.stdout[-2000:] + ref.stderr[-2000:]
    r = run("sp", *args)
    assert r.returncode == 0, r.stdout + r.stderr
    missing = _report_keys(ref.stdout) - _report_keys(r.stdout)
    assert not missing, sorted(missing)
    # the values parse.py compares across runs agree where they describe the problem
    for key in ("Number of time steps", "Time unroll", "Intra-diamond width", "Stencil Kernel semi-bandwidth",
                "Precision", "Multi-wavefront updates", "Wavefront width"):
        ref_v = re.search(r"^%s:\s*(.*)$" % re.escape(key), ref.stdout, re.M)
        gpu_v = re.search(r"^%s:\s*(.*)$" % re.escape(key), r.stdout, re.M)
        if ref_v:
            assert gpu_v and gpu_v.group(1).strip() == ref_v.group(1).strip(), key
