"""girih's own driver with the MI355X stepper registered in its TSList.

oracle/Makefile (target gpu) links girih_amd/integration/gpu_ts.c -- the stepper a
girih maintainer adds, compiled against girih's own src/data_structures.h -- and a
main that defines TSList with it, with girih's objects built from its sources.
Everything around the GPU run is then girih's code: parse_args, init (including the
diamond stepper's validity rules and nt rounding), the closed-form initial state,
verify() with its serial reference and exact compare (verification.c:26-51,
823-860), and performance_test() with its report.  The binaries live under
oracle/_ref (built where girih's sources are present); these tests skip without
them.
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(REPO, "oracle", "_ref")
PASS = re.compile(r"eMax:0\.000e\+00\|eL1:0\.000e\+00-PASSED")

pytestmark = pytest.mark.gpu


def _run(name, args, timeout=300):
    exe = os.path.join(REF, name)
    if not os.path.exists(exe):
        pytest.skip("%s not built (make -C oracle gpu, where girih's sources are present)" % name)
    print(name, " ".join(str(a) for a in args), flush=True)
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout,
                          env=dict(os.environ, OMP_NUM_THREADS="1"))  # girih checks its thread groups against the diamonds


DIAMOND = ["--thread-group-size", 1, "--thx", 1, "--thy", 1, "--thz", 1, "--num-wavefronts", 1]


@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("kernel,shape,nt", [(0, (40, 24, 20), 9), (1, (70, 33, 21), 10), (5, (37, 20, 18), 7)])
def test_girih_verify_passes_with_gpu_stepper(prec, kernel, shape, nt):
    """TSList[3] = "MI355X" (spatial-blocking semantics) under girih's verify()."""
    nx, ny, nz = shape
    r = _run("mwd_kernel_gputs_" + prec, ["--target-ts", 3, "--target-kernel", kernel, "--nx", nx, "--ny", ny,
                                         "--nz", nz, "--nt", nt, "--verify", 1])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "MI355X" in r.stdout
    assert PASS.search(r.stdout), r.stdout[-2000:]


@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("kernel,shape,nt,t_dim", [(0, (40, 32, 24), 9, 1), (1, (70, 32, 24), 9, 3),
                                                   (5, (33, 64, 30), 11, 7)])
def test_girih_verify_passes_with_gpu_diamond_stepper(prec, kernel, shape, nt, t_dim):
    """TSList[2] = "Diamond" run on the GPU: girih's diamond initialisation rounds nt
    and checks the diamond width first; girih's verify() compares exactly."""
    nx, ny, nz = shape
    r = _run("mwd_kernel_gpudiam_" + prec, ["--target-ts", 2, "--target-kernel", kernel, "--nx", nx, "--ny", ny,
                                           "--nz", nz, "--nt", nt, "--t-dim", t_dim, "--verify", 1] + DIAMOND)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert PASS.search(r.stdout), r.stdout[-2000:]


def test_girih_performance_test_times_gpu_stepper():
    """girih's performance_test() (n_tests, median, report) around the GPU stepper."""
    r = _run("mwd_kernel_gputs_sp", ["--target-ts", 3, "--target-kernel", 0, "--nx", 128, "--ny", 128, "--nz", 128,
                                     "--nt", 20, "--n-tests", 3])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Time stepper name: MI355X" in r.stdout
    m = re.search(r"RANK0 GStencil/s MEDIAN: ([0-9.]+)", r.stdout)
    assert m and float(m.group(1)) > 0
    assert "COMPLETED SUCCESSFULLY" in r.stdout


def test_girih_rejects_bad_diamond_width_before_gpu():
    """girih's own validity check runs before the GPU stepper is reached."""
    r = _run("mwd_kernel_gpudiam_dp", ["--target-ts", 2, "--target-kernel", 0, "--nx", 16, "--ny", 24, "--nz", 40,
                                       "--nt", 9, "--t-dim", 1, "--verify", 1] + DIAMOND)
    assert r.returncode != 0
    assert "diamond width" in (r.stdout + r.stderr)


# girih's regression matrix (scripts/verification), one rank: verification_std.py's
# grids x kernels 0-5 for the spatial-blocking semantics, and the grids and diamond
# depths of verification_idiam.py's general part for the diamond semantics, in both
# precisions, with girih's nt = max(10, nx/R/2) (verification_utils.py:14).
RADIUS = {0: 4, 1: 1, 2: 1, 3: 1, 4: 4, 5: 1}
STD_GRIDS = [(16, 32, 64), (32, 64, 16), (64, 16, 32)]
IDIAM_GRIDS = [(16, 32, 128, 1), (16, 64, 128, 3), (16, 64, 128, 7), (32, 128, 128, 7)]


def _nt(nx, kernel):
    return max(10, nx // RADIUS[kernel] // 2)


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5])
def test_girih_verification_std_matrix(prec, kernel):
    for nx, ny, nz in STD_GRIDS:
        r = _run("mwd_kernel_gputs_" + prec, ["--target-ts", 3, "--target-kernel", kernel, "--nx", nx, "--ny", ny,
                                             "--nz", nz, "--nt", _nt(nx, kernel), "--verbose", 0, "--verify", 1])
        assert r.returncode == 0 and PASS.search(r.stdout), ((nx, ny, nz), r.stdout[-1500:] + r.stderr[-1500:])


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5])
def test_girih_verification_idiam_matrix(prec, kernel):
    for nx, ny, nz, t_dim in IDIAM_GRIDS:
        r = _run("mwd_kernel_gpudiam_" + prec, ["--target-ts", 2, "--target-kernel", kernel, "--nx", nx, "--ny", ny,
                                               "--nz", nz, "--nt", _nt(nx, kernel), "--t-dim", t_dim, "--verbose", 0,
                                               "--verify", 1, "--thread-group-size", 1, "--thx", 1, "--thy", 1,
                                               "--thz", 1, "--num-wavefronts", 4, "--mwd-type", 1])
        assert r.returncode == 0 and PASS.search(r.stdout), ((nx, ny, nz, t_dim), r.stdout[-1500:] + r.stderr[-1500:])
