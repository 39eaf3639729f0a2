"""Benchmark: 25-point wave stencil (girih kernel 0) at 1024^3 on one MI355X.

Workload: u^(n+1) = 2u^n - u^(n-1) + (v dt/h)^2 Lap8(u^n), 8th-order FD weights,
v = 3000 m/s, h = 10 m, dt at CFL 0.4 (roc2 = 0.16), 4-cell halo, 15 Hz Ricker
point source at the grid centre, zero initial fields; one "step" = one time step
over the whole 1024^3 grid.  The timed run is the FP32 path (FP32 fields, neighbour
pairs summed before scaling, compensated 2u^n - u^(n-1) update); the same steps are
then run in FP64 (the GPU oracle) and the FP32 error is reduced on the device in FP64.

Prints ONE JSON line (rank 0, N = 1).  --gpus N > 1 is refused: this build runs on
one GPU.

    python bench.py [--steps 1000] [--warmup 10] [--n 1024] [--no-fp64] [--no-cpu-baseline]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def _sync():
    import torch
    torch.cuda.synchronize()


def run_wave(dtype, flags, cfg, steps, warmup):
    from girih_amd import gpu
    st = gpu.Stepper(0, cfg["nx"], cfg["ny"], cfg["nz"], dtype=dtype, flags=flags | gpu.GR_ROC2_SCALAR,
                     roc2=cfg["roc2"], coef=cfg["coef"])
    tuned = st.autotune(4)  # tile shape x z-chunk search on scratch copies of the fields
    if warmup > 0:
        st.init_fields(2)
        st.step(warmup)
    st.init_fields(2)
    st.set_source(*cfg["src_point"], cfg["src"][:steps + 2])
    _sync()
    t0 = time.perf_counter()
    st.step(steps)
    _sync()
    t1 = time.perf_counter()
    stats = st.stats()
    stats["tuned_ms_per_step"] = tuned["last_ms"]
    return st, t1 - t0, stats


def cpu_baseline(n, budget_nt):
    """girih's own OpenMP diamond stepper (1WD: one thread per diamond, t_dim 1),
    built from its sources into oracle/_ref, on the host cores of this box."""
    exe = os.path.join(REPO, "oracle", "_ref", "mwd_kernel_dp")
    cores = min(int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1, 16)
    if not os.path.exists(exe):
        return cpu_baseline_port()
    nt = budget_nt
    cmd = [exe, "--nx", str(n), "--ny", str(n), "--nz", str(n), "--nt", str(nt), "--target-ts", "2",
           "--target-kernel", "0", "--thread-group-size", "1", "--thx", "1", "--thy", "1", "--thz", "1",
           "--t-dim", "1", "--num-wavefronts", "1", "--n-tests", "1", "--verbose", "0"]
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="true", OMP_PLACES="cores",
               OMP_WAIT_POLICY="active")
    t0 = time.perf_counter()
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    wall = time.perf_counter() - t0
    m = re.search(r"Total RANK0 MStencil/s MAX:\s*([0-9.eE+-]+)", r.stdout)
    if r.returncode != 0 or not m:
        return {"value": None, "unit": "GLUP/s", "cores": cores, "kind": "reference",
                "sample": "failed: rc %d %s" % (r.returncode, (r.stderr or r.stdout)[-200:])}
    return {"value": float(m.group(1)) / 1e3, "unit": "GLUP/s", "cores": cores, "kind": "reference",
            "sample": "girih mwd_kernel (DP, OpenMP 1WD diamond, t_dim 1, kernel 0 with girih's own "
                      "coefficients) %d^3 x nt %d, n-tests 1, %.1f s wall incl. init" % (n, nt, wall)}


def cpu_baseline_port():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import oracle as orc
    n, nt = 128, 4
    u, v, roc2 = orc.init_ref(0, n, n, n)
    coef = orc.init_coef(0, u.shape)
    t0 = time.perf_counter()
    orc.run(0, coef, u, v, roc2, nt=nt)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * nt / dt / 1e9, "unit": "GLUP/s", "cores": 1, "kind": "port",
            "sample": "serial CPU oracle (-O0), 25-pt wave FP64 %d^3 x %d steps" % (n, nt)}


def load_traffic(key):
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-nt", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the z-slab sharding across GPUs is designed, not built)" % args.gpus, file=sys.stderr)
        return 2

    import torch
    from girih_amd import gpu, physics
    torch.cuda.set_device(0)
    n, K, W = args.n, args.steps, args.warmup
    cfg = physics.wave_1024_config(nx=n, ny=n, nz=n, nt=K)
    cells = n * n * n

    st32, t32, s32 = run_wave(gpu.GR_F32, gpu.GR_FAST, cfg, K, W)
    glups32 = cells * K / t32 / 1e9
    launch_ms = s32["last_ms"] / max(1, s32["last_launches"])
    steps_per_launch = K / max(1, s32["last_launches"])
    # algorithmic HBM bytes per launch: read u^n, u^(n-1), write u^(n+1) per cell and
    # fused step (4 B each, roc2 is a scalar); a fused launch streams 2 fields in, 2 out
    bytes_per_cell_launch = 12.0 if steps_per_launch == 1 else 16.0
    algo_bytes = bytes_per_cell_launch * cells
    achieved = algo_bytes / (launch_ms * 1e-3) / 1e9
    traffic = load_traffic("wave25_f32_%d" % n)

    out = {
        "metric": "GLUP/s + % HBM roofline, 25pt wave 1024^3; FP32 vs FP64 GPU at stated error",
        "value": round(glups32, 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(t32 * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (zero initial field, Ricker point source)",
        "config": {"workload": "25-pt star wave (girih kernel 0), %d^3 + 4-cell halo, %d steps, v=3000 m/s, "
                               "h=10 m, CFL 0.4, 15 Hz Ricker at centre, FP32 fields, compensated update" % (n, K),
                   "nx": n, "ny": n, "nz": n, "nt": K, "t_block": s32["t_block"],
                   "tile": [s32["tile_x"], s32["tile_y"]], "zchunk": s32["zchunk"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algo_bytes_per_launch": algo_bytes, "avg_launch_ms": round(launch_ms, 5),
                     "steps_per_launch": steps_per_launch},
    }
    if not args.no_fp64:
        st64, t64, s64 = run_wave(gpu.GR_F64, 0, cfg, K, W)
        glups64 = cells * K / t64 / 1e9
        field = (K - 1) % 2  # after K applications the last written field is U1 (odd) / U2 (even)
        cmp = st32.compare(field, st64, field)
        out["fp64"] = {"glups": round(glups64, 3), "ms_per_step": round(t64 * 1e3 / K, 4),
                       "fp32_over_fp64": round(glups32 / glups64, 3),
                       "rel_l2_fp32_vs_fp64": cmp["rel_l2"], "max_abs_err": cmp["max_diff"],
                       "max_abs_ref": cmp["max_ref"]}
        st64.close()
    st32.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_nt)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
