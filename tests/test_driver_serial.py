"""The standalone driver's --verify reference, pinned directly against the oracle.

`mwd_kernel_gpu --verify 1` compares the GPU run with a serial evaluation of its
own (serial_reference in girih_amd/driver/mwd_kernel_gpu.c: girih's closed-form
initial state rebuilt by the driver's domain_data_fill, the driver's coefficients,
and serial sweeps in stencils.c order).  Here that code runs on the CPU, built
from the driver's source into a small harness, and its U1/U2 must equal the
oracle's (verification.c restated, pinned against girih's own binaries) bit for
bit, for kernels 0-5, DP and SP, the step counts of girih's SB stepper (ts 0) and
of its diamond stepper (ts 2).  No GPU call is made.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "girih_amd", "driver")
LIB = os.path.join(REPO, "girih_amd", "lib")

HARNESS = r"""
#define main gr_driver_main
#include "mwd_kernel_gpu.c"
#undef main
/* argv: the driver's flags, then the output path.  Writes ux, uy, uz (int32) and
 * the serial reference's U1 and U2 (real_t). */
int main(int argc, char **argv) {
  const char *out = argv[argc - 1];
  Params p;
  param_default(&p);
  parse_args(argc - 1, argv, &p);
  init(&p);
  arrays_allocate(&p);
  init_coeff(&p);
  domain_data_fill(&p);
  /* the step count gpu_ts records from gr_run_ts (girih's rounding for ts 2; host-only) */
  p.nt_applied = p.target_ts == 2 ? gr_mwd_round_nt(p.nt, p.t_dim) : p.nt;
  real_t *u = NULL, *v = NULL;
  serial_reference(&p, &u, &v);
  const int dims[3] = {p.nx + 2 * p.R, p.ld[1], p.ld[2]};
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  FILE *f = fopen(out, "wb");
  if (!f) return 2;
  fwrite(dims, sizeof(int), 3, f);
  fwrite(u, sizeof(real_t), n, f);
  fwrite(v, sizeof(real_t), n, f);
  fclose(f);
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "libgirih_gpu.so")):
        pytest.skip("libgirih_gpu.so not built")
    d = tmp_path_factory.mktemp("serial")
    src = d / "harness.c"
    src.write_text(HARNESS)
    exes = {}
    for prec, dp in (("dp", 1), ("sp", 0)):
        exe = str(d / ("harness_" + prec))
        subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-DDP=%d" % dp, "-I" + os.path.join(REPO, "include"),
                        "-I" + DRIVER, str(src), "-o", exe, "-L" + LIB, "-lgirih_gpu", "-Wl,-rpath," + LIB, "-lm"],
                       check=True, capture_output=True)
        exes[prec] = exe
    return exes, d


CASES = [  # kernel, nx, ny, nz, nt, ts, extra flags
    (0, 16, 24, 20, 5, 0, []),
    (1, 20, 18, 22, 9, 0, []),
    (2, 12, 14, 16, 4, 0, []),
    (3, 12, 14, 16, 4, 0, []),
    (4, 16, 18, 20, 3, 0, []),
    (5, 20, 18, 22, 7, 0, []),
    (0, 16, 64, 40, 12, 2, ["--t-dim", 1]),
    (1, 16, 32, 24, 10, 2, ["--t-dim", 3]),
]


@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "k%d_ts%d" % (c[0], c[5]))
def test_driver_serial_reference_equals_oracle(harness, prec, case):
    exes, d = harness
    kernel, nx, ny, nz, nt, ts, extra = case
    out = d / ("out_%s_k%d_ts%d" % (prec, kernel, ts))
    args = [exes[prec], "--target-kernel", kernel, "--target-ts", ts, "--nx", nx, "--ny", ny, "--nz", nz, "--nt", nt,
            "--thread-group-size", 1, *extra, str(out)]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dt = np.float64 if prec == "dp" else np.float32
    raw = out.read_bytes()
    ux, uy, uz = np.frombuffer(raw[:12], dtype=np.int32)
    n = int(ux) * int(uy) * int(uz)
    body = np.frombuffer(raw[12:], dtype=dt)
    U1 = body[:n].reshape(uz, uy, ux)
    U2 = body[n:2 * n].reshape(uz, uy, ux)
    u, v, roc2 = orc.init_ref(kernel, nx, ny, nz, dtype=dt)
    assert (uz, uy, ux) == u.shape
    coef = orc.init_coef(kernel, u.shape, dtype=dt)
    if ts == 2:  # girih's diamond stepper: nt rounded, nt - 1 applications, U1 written first
        napply = orc.mwd_round_nt(nt, extra[1]) - 1
        for a in range(napply):
            if a % 2 == 0:
                orc.sweep(kernel, coef, u, v, roc2)
            else:
                orc.sweep(kernel, coef, v, u, roc2)
    else:  # 2*ceil(nt/2) applications
        orc.run(kernel, coef, u, v, roc2, nt=nt)
    np.testing.assert_array_equal(U1, u)
    np.testing.assert_array_equal(U2, v)
