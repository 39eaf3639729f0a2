"""Build the HIP library (gfx950) and the C driver in-tree.

    python -m girih_amd.build            # libgirih_gpu.so + mwd_kernel_gpu
    python -m girih_amd.build --lib      # library only

Outputs: girih_amd/lib/libgirih_gpu.so and girih_amd/driver/bin/mwd_kernel_gpu_{dp,sp}
(the FP64 and FP32 builds of the driver, like girih's `make dp` / `make`).
"""
import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libgirih_gpu.so")
DRIVER_DIR = os.path.join(PKG, "driver")
DRIVER_BINS = {"dp": os.path.join(DRIVER_DIR, "bin", "mwd_kernel_gpu_dp"),
               "sp": os.path.join(DRIVER_DIR, "bin", "mwd_kernel_gpu_sp")}

HIP_SOURCES = ["gr_launch_f32.hip", "gr_launch_f64.hip", "gr_launch_bf16.hip", "gr_runtime.hip", "gr_tblock.hip",
               "gr_wave2.hip", "gr_fused1.hip", "gr_model.hip", "gr_batch.hip", "gr_adjoint.hip"]
ARCH = os.environ.get("GIRIH_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the reference-order kernels must not fuse multiply-adds
# (bit-exact FP64/FP32 against girih's serial reference); FAST kernels call fma
# explicitly.  packed-fp32-ops is disabled for all device code.
HIPCC_FLAGS = ["-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument", "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(verbose=False):
    """Compile every HIP source (in parallel, MAX_JOBS or up to 8 at once) and link the library."""
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        jobs.append(([hipcc()] + HIPCC_FLAGS + ["-I" + os.path.join(REPO, "include"), "-c", path, "-o", obj], obj))
    width = max(1, min(8, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    running, objs = [], []
    for cmd, obj in jobs:
        print("+", " ".join(cmd), flush=True)
        running.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
        while len(running) >= width:
            _wait_first(running)
    while running:
        _wait_first(running)
    tmp = LIB + ".tmp"
    run([hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def _wait_first(running):
    proc, cmd = running.pop(0)
    if proc.wait() != 0:
        for p, _ in running:
            p.wait()
        raise subprocess.CalledProcessError(proc.returncode, cmd)


def build_driver():
    src = os.path.join(DRIVER_DIR, "mwd_kernel_gpu.c")
    os.makedirs(os.path.join(DRIVER_DIR, "bin"), exist_ok=True)
    for prec, out in DRIVER_BINS.items():
        # -ffp-contract=off: the --verify serial check keeps girih's operand order exactly
        run(["gcc", "-O2", "-std=c11", "-Wall", "-ffp-contract=off", "-DDP=%d" % (prec == "dp"),
             "-I" + os.path.join(REPO, "include"), "-o", out, src, "-L" + LIBDIR, "-lgirih_gpu",
             "-Wl,-rpath,$ORIGIN/../../lib", "-lm"])
    return DRIVER_BINS


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    args = ap.parse_args(argv)
    build_lib()
    if not args.lib:
        build_driver()


if __name__ == "__main__":
    sys.exit(main())
