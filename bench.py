"""Benchmark: 25-point wave stencil (girih kernel 0) at 1024^3 on one MI355X.

Workload: u^(n+1) = 2u^n - u^(n-1) + (v dt/h)^2 Lap8(u^n), 8th-order FD weights,
v = 3000 m/s, h = 10 m, dt at CFL 0.4 (roc2 = 0.16), 4-cell halo, 15 Hz Ricker
point source at the grid centre, zero initial fields; one "step" = one time step
over the whole 1024^3 grid.  The timed run is the FP32 path (FP32 fields,
neighbour pairs summed before scaling, increment-form leapfrog update); the same
steps are then run in FP64 (bit-identical to girih's serial reference) and the
FP32 error is reduced on the device in FP64.

Prints ONE JSON line (N = 1).  --gpus N > 1 is refused: this build runs on one GPU.

    python bench.py [--steps 1000] [--warmup 10] [--n 1024] [--no-fp64] [--no-cpu-baseline]
                    [--no-extras] [--config VARIANT,ZCHUNK,XCD_ORDER,T_BLOCK]

--config pins the launch configuration instead of autotuning it (profiler runs
use the configuration the tuned run printed).  The "extras" object measures the
other workloads of the same stepper: the 7-point Jacobi at 128^3 (FP64), the
7-point variable-coefficient stencil at 512^3 (FP64 and FP32, with the FP32
error), the batch of 16 uneven 25-point problems with bf16 roc2, and the adjoint
gradient of a 400^3 x 2000-step shot (FP32, checkpointed) with its FP64 and
finite-difference checks.
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


def _configure(st, config, tune_steps=4):
    """Autotune (tile variant x z chunk x tile order x fused depth) or pin a config."""
    if config is None:
        return st.autotune(tune_steps)
    var, zc, order, tb = config
    st.set_config(var, zc, tb)
    st.set_xcd_order(order)
    return st.stats()


def run_wave(dtype, flags, cfg, steps, warmup, config=None):
    from girih_amd import gpu
    st = gpu.Stepper(0, cfg["nx"], cfg["ny"], cfg["nz"], dtype=dtype, flags=flags | gpu.GR_ROC2_SCALAR,
                     roc2=cfg["roc2"], coef=cfg["coef"])
    tuned = _configure(st, config)
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
    try:
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    except (OSError, subprocess.SubprocessError) as e:
        return {"value": None, "unit": "GLUP/s", "cores": cores, "kind": "reference", "sample": "failed: %s" % e}
    wall = time.perf_counter() - t0
    m = re.search(r"Total RANK0 MStencil/s MAX:\s*([0-9.eE+-]+)", r.stdout)
    if r.returncode != 0 or not m:
        return {"value": None, "unit": "GLUP/s", "cores": cores, "kind": "reference",
                "sample": "failed: rc %d %s" % (r.returncode, (r.stderr or r.stdout)[-200:])}
    out = {"value": float(m.group(1)) / 1e3, "unit": "GLUP/s", "cores": cores, "kind": "reference",
           "sample": "girih mwd_kernel (DP, OpenMP 1WD diamond, t_dim 1, kernel 0 with girih's own "
                     "coefficients) %d^3 x nt %d, n-tests 1, %.1f s wall incl. init" % (n, nt, wall)}
    # girih's spatially blocked stepper (ts 0) on the same cores, reported beside the diamond stepper
    sb_nt = 6
    cmd_sb = [exe, "--nx", str(n), "--ny", str(n), "--nz", str(n), "--nt", str(sb_nt), "--target-ts", "0",
              "--target-kernel", "0", "--thread-group-size", "1", "--thx", "1", "--thy", "1", "--thz", "1",
              "--t-dim", "1", "--num-wavefronts", "1", "--n-tests", "1", "--verbose", "0"]
    try:
        r = subprocess.run(cmd_sb, env=env, capture_output=True, text=True, timeout=300)
        m = re.search(r"RANK0 GStencil/s MEDIAN:\s*([0-9.eE+-]+)", r.stdout)
        out["sb_value"] = float(m.group(1)) if (r.returncode == 0 and m) else None
    except (OSError, subprocess.SubprocessError):
        out["sb_value"] = None
    out["sb_sample"] = "girih mwd_kernel spatial blocking (ts 0, DP) %d^3 x nt %d, n-tests 1" % (n, sb_nt)
    return out


def cpu_baseline_port():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
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


def _timed_steps(st, napply):
    _sync()
    t0 = time.perf_counter()
    st.step(napply)
    _sync()
    return time.perf_counter() - t0


def extra_first_order(kernel, n, nt, coef=None, fp32=True, graph=False, bf16=False):
    """girih kernels 1 / 5: FP64 (bit-exact path) and optionally FP32 on the same
    seeded input, autotuned; FP32 error vs FP64 reduced on the device.  graph=True
    replays the step loop from a HIP graph (captured by an untimed first run)."""
    from girih_amd import gpu
    napply = 2 * ((nt + 1) // 2)  # girih's SB/MWD application count for even nt
    out = {}
    runs = [("f64", gpu.GR_F64, 0)] + ([("f32", gpu.GR_F32, 0)] if fp32 else [])
    if bf16:  # FP32 fields, coefficient arrays stored as bf16 (same seed, rounded to nearest even)
        runs.append(("f32_bf16coef", gpu.GR_F32, gpu.GR_COEF_BF16))
    keep = {}
    for name, dtype, flags in runs:
        st = gpu.Stepper(kernel, n, n, n, dtype=dtype, flags=flags, coef=coef)
        if gpu.NUM_COEF_ARRAYS[kernel]:
            st.init_coef_uniform(5, 0.3, 0.5, 0.0, 0.1)
        st.init_fields(1, 12345)
        tuned = st.autotune(4)
        if graph:
            st.set_graph(True)
            st.init_fields(1, 12345)
            st.step(napply)  # capture
        st.init_fields(1, 12345)
        t = _timed_steps(st, napply)
        s = st.stats()
        b_sweep = 2 * (8 if dtype == gpu.GR_F64 else 4) + gpu.NUM_COEF_ARRAYS[kernel] * (
            2 if flags & gpu.GR_COEF_BF16 else (8 if dtype == gpu.GR_F64 else 4))
        gl = n ** 3 * napply / t / 1e9
        out[name] = {"glups": round(gl, 2), "ms_per_step": round(t * 1e3 / napply, 4),
                     "sweep_roofline_frac": round(gl * b_sweep / HBM_PEAK_GBS, 4), "bytes_per_lup_sweep": b_sweep,
                     "t_block": tuned["t_block"], "zchunk": tuned["zchunk"], "tile": [tuned["tile_x"],
                                                                                     tuned["tile_y"]],
                     "launches": s["last_launches"], "graph": graph}
        keep[name] = st
    lowp = [k for k in keep if k != "f64"]
    for k in lowp:
        c = keep[k].compare(1, keep["f64"], 1)
        out[k]["rel_l2_vs_f64"] = c["rel_l2"]
        out[k]["f64_max_abs_after_run"] = c["max_ref"]  # below FP32's range when the stencil contracts
    if lowp:
        for st in keep.values():  # the error while the field is still representable in FP32
            st.init_fields(1, 12345)
            st.step(100)
        for k in lowp:
            out[k]["rel_l2_vs_f64_after_100"] = keep[k].compare(1, keep["f64"], 1)["rel_l2"]
        if "f32_bf16coef" in out:
            out["f32_bf16coef"]["note"] = "error includes the bf16 rounding of the coefficients"
    for st in keep.values():
        st.close()
    return out


def extra_batch(nt, seed=2024):
    """16 independent 25-pt problems, dims uniform in 192..512 (multiples of 16),
    8-layer velocity model with +-5% Gaussian perturbation, bf16 roc2, one
    persistent work-queue launch per step replayed from a HIP graph; each problem
    checked against the FP64 path (bit-identical to the CPU oracle) on the same
    bf16 roc2 values."""
    import numpy as np
    from girih_amd import gpu, physics
    rng = np.random.default_rng(seed)
    sizes = [tuple(int(x) for x in rng.choice(np.arange(192, 513, 16), 3)) for _ in range(16)]
    vmax = 4500.0 * 1.05
    h = 10.0
    dt = physics.cfl_dt(vmax, h, 0.4)
    f0 = 15.0
    src = physics.roc2_const(3000.0, dt, h) * physics.ricker(f0, dt, nt + 2)
    cells = sum(a * b * c for a, b, c in sizes)
    with gpu.Batch(sizes, flags=gpu.GR_FAST | gpu.GR_COEF_BF16, coef=physics.fd8_coef()) as b:
        for i, p in enumerate(b.problems):
            p.init_fields(2)
            p.init_roc2_model(gpu.vmodel(seed + i, dt / h))
            nx, ny, nz = sizes[i]
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
        var, zc = b.autotune(4, use_graph=True)
        b.step(nt, use_graph=True)  # captures the step-loop graph, then runs it
        t_first = b.stats()["last_ms"]
        for i, p in enumerate(b.problems):  # restart from the initial state, replay the captured graph
            p.init_fields(2)
            nx, ny, nz = sizes[i]
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
        b.step(nt, use_graph=True)
        s = b.stats()
        errs = []
        for i, p in enumerate(b.problems):
            nx, ny, nz = sizes[i]
            with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F64, coef=physics.fd8_coef()) as ref:
                ref.init_fields(2)
                ref.init_roc2_model(gpu.vmodel(seed + i, dt / h, round_bf16=True))
                ref.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
                ref.set_config(zchunk=64)
                ref.step(nt)
                f = (nt - 1) % 2
                errs.append(p.compare(f, ref, f)["rel_l2"])
    gl = cells * nt / (s["last_ms"] * 1e-3) / 1e9
    return {"glups": round(gl, 2), "ms_per_step": round(s["last_ms"] / nt, 4), "problems": len(sizes),
            "cells": cells, "steps": nt, "sizes": sizes, "items_per_step": s["items"],
            "persistent_workgroups": s["grid"], "tile": [s["tile_x"], s["tile_y"]], "zchunk": s["zchunk"],
            "graph_first_call_ms": round(t_first, 2),
            "sweep_roofline_frac": round(gl * 14 / HBM_PEAK_GBS, 4), "bytes_per_lup_sweep": 14,
            "rel_l2_vs_f64_max": max(errs), "rel_l2_vs_f64": [float("%.3g" % e) for e in errs]}


def extra_adjoint(n=400, nt=2000, ckpt=50, nfd=20, seed=7):
    """Adjoint gradient of the 25-pt variable-velocity stepper: n^3, nt steps,
    h = 10 m, 8-layer velocity (+-5% blobs), 15 Hz Ricker at 50 m depth, receiver
    plane at z = 8 cells every step, misfit 0.5 ||d_pred - d_obs||^2 with d_obs from
    a +3% Gaussian-blob velocity perturbation; FP32 shot with checkpoints every
    `ckpt` steps through torch.autograd, checked against the FP64 GPU adjoint
    (bit-identical to the FP64 CPU adjoint at test sizes) and against central
    finite differences of the FP64 forward model on `nfd` cells the wave reached."""
    import numpy as np
    import torch
    from girih_amd import adjoint, gpu, physics
    h = 10.0
    dt = physics.cfl_dt(4500.0 * 1.05, h, 0.4)
    dth = dt / h
    src = physics.roc2_const(1500.0, dt, h) * physics.ricker(15.0, dt, nt)
    src_pt = (4 + n // 2, 4 + n // 2, 4 + 5)
    rec_k = 4 + 8

    def slowness2(anomaly):
        with gpu.Stepper(0, n, n, n, dtype=gpu.GR_F64, coef=physics.fd8_coef()) as st:
            st.init_roc2_model(gpu.vmodel(seed, dth, anomaly=anomaly, anomaly_center=(n / 2, n / 2, n / 2),
                                          anomaly_sigma=n / 10.0))
            r2 = st.download_roc2()[4:-4, 4:-4, 4:4 + n]
        return np.ascontiguousarray(dth ** 2 / r2)

    m_obs, m0 = slowness2(0.03), slowness2(0.0)
    s64 = adjoint.Shot(n, n, n, nt, dth, src_pt, src, rec_k, ckpt=ckpt, dtype=gpu.GR_F64)
    s32 = adjoint.Shot(n, n, n, nt, dth, src_pt, src.astype(np.float32).astype(np.float64), rec_k, ckpt=ckpt,
                       dtype=gpu.GR_F32)
    try:
        d_obs = s64.forward(torch.tensor(m_obs, device="cuda"))
        m64 = torch.tensor(m0, device="cuda")
        tr64 = s64.forward(m64)
        g64 = s64.vjp(m64, tr64 - d_obs)
        # FP32: forward + backward through the autograd Function, timed
        m32 = torch.tensor(m0, dtype=torch.float32, device="cuda", requires_grad=True)
        _sync()
        t0 = time.perf_counter()
        tr32 = adjoint.shot_traces(m32, s32)
        _sync()
        t1 = time.perf_counter()
        loss = 0.5 * ((tr32.double() - d_obs) ** 2).sum()
        loss.backward()
        _sync()
        t2 = time.perf_counter()
        g32 = m32.grad.double()
        rel = float(torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64))
        J0 = float(0.5 * ((tr64 - d_obs) ** 2).sum())
        # the FP32 gradient error follows from the FP32 traces' error relative to the residual
        tr_err = float(torch.linalg.norm(tr32.double() - tr64) / torch.linalg.norm(tr64))
        res_rel = float(torch.linalg.norm(tr64 - d_obs) / torch.linalg.norm(tr64))
        del tr32, tr64
        g = g64.cpu().numpy()
        g32h = g32.cpu().numpy()
        gmax = np.abs(g).max()
        live = np.argwhere(np.abs(g) > 1e-2 * gmax)
        rng = np.random.default_rng(seed)
        errs32, errs64, errs32n = [], [], []

        def J(mm):
            t = s64.forward(torch.tensor(mm, device="cuda"))
            return 0.5 * float(((t - d_obs) ** 2).sum())

        for q in rng.choice(len(live), min(nfd, len(live)), replace=False):
            k, j, i = live[q]
            hh = 1e-4 * m0[k, j, i]
            mp, mm = m0.copy(), m0.copy()
            mp[k, j, i] += hh
            mm[k, j, i] -= hh
            fd = (J(mp) - J(mm)) / (2 * hh)
            errs32.append(abs(g32h[k, j, i] - fd) / abs(fd))
            errs64.append(abs(g[k, j, i] - fd) / abs(fd))
            errs32n.append(abs(g32h[k, j, i] - fd) / gmax)
    finally:
        s32.close()
        s64.close()
    fwd_glups = n ** 3 * nt / (t1 - t0) / 1e9
    return {"grid": n, "steps": nt, "checkpoint_every": ckpt, "forward_s": round(t1 - t0, 3),
            "backward_s": round(t2 - t1, 3), "forward_glups": round(fwd_glups, 2),
            "backward_over_forward": round((t2 - t1) / (t1 - t0), 3), "misfit": J0,
            "rel_l2_grad_fp32_vs_fp64": rel, "fd_cells": len(errs32),
            "fd_max_rel_err_fp32_adjoint": max(errs32), "fd_max_rel_err_fp64_adjoint": max(errs64),
            "fd_max_err_fp32_adjoint_over_max_grad": max(errs32n),
            "rel_l2_traces_fp32_vs_fp64": tr_err, "residual_over_traces": res_rel}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-nt", type=int, default=60)
    ap.add_argument("--config", default=None, help="VARIANT,ZCHUNK,XCD_ORDER,T_BLOCK (skips the autotuner)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the z-slab sharding across GPUs is designed, not built)" % args.gpus, file=sys.stderr)
        return 2
    config = tuple(int(x) for x in args.config.split(",")) if args.config else None
    if config is not None and len(config) != 4:
        print("bench.py: --config needs VARIANT,ZCHUNK,XCD_ORDER,T_BLOCK", file=sys.stderr)
        return 2

    import torch
    from girih_amd import gpu, physics
    torch.cuda.set_device(0)
    n, K, W = args.n, args.steps, args.warmup
    cfg = physics.wave_1024_config(nx=n, ny=n, nz=n, nt=K)
    cells = n * n * n

    st32, t32, s32 = run_wave(gpu.GR_F32, gpu.GR_FAST, cfg, K, W, config)
    glups32 = cells * K / t32 / 1e9
    launch_ms = s32["last_ms"] / max(1, s32["last_launches"])
    steps_per_launch = K / max(1, s32["last_launches"])
    # algorithmic HBM bytes per launch: read u^n, u^(n-1), write u^(n+1) per cell and
    # step (4 B each, roc2 is a scalar); a fused launch streams 2 fields in, 2 out
    bytes_per_cell_launch = 12.0 if steps_per_launch == 1 else 16.0
    algo_bytes = bytes_per_cell_launch * cells
    achieved = algo_bytes / (launch_ms * 1e-3) / 1e9
    launch_config = "%d,%d,%d,%d" % (s32["tile_variant"], s32["zchunk"], s32["xcd_order"], s32["t_block"])
    # profiled traffic is kept per launch configuration (it depends on the tile and z chunk)
    tr = load_traffic("wave25_f32_%d@%s" % (n, launch_config))
    traffic = tr["bytes_per_launch"] if tr and tr.get("launch_config") == launch_config else None

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
                               "h=10 m, CFL 0.4, 15 Hz Ricker at centre, FP32 fields, increment-form update"
                               % (n, K),
                   "nx": n, "ny": n, "nz": n, "nt": K, "t_block": s32["t_block"],
                   "tile": [s32["tile_x"], s32["tile_y"]], "zchunk": s32["zchunk"],
                   "launch_config": launch_config,
                   "autotuned": config is None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algo_bytes_per_launch": algo_bytes, "avg_launch_ms": round(launch_ms, 5),
                     "steps_per_launch": steps_per_launch,
                     "traffic_bytes_per_lup": (round(traffic / (cells * steps_per_launch), 3) if traffic else None),
                     "traffic_profile": tr.get("summary") if traffic else None},
    }
    if not args.no_fp64:
        st64, t64, s64 = run_wave(gpu.GR_F64, 0, cfg, K, W)
        glups64 = cells * K / t64 / 1e9
        field = (K - 1) % 2  # after K applications the last written field is U1 (odd K) / U2 (even K)
        cmp = st32.compare(field, st64, field)
        out["fp64"] = {"glups": round(glups64, 3), "ms_per_step": round(t64 * 1e3 / K, 4),
                       "fp32_over_fp64": round(glups32 / glups64, 3),
                       "rel_l2_fp32_vs_fp64": cmp["rel_l2"], "max_abs_err": cmp["max_diff"],
                       "max_abs_ref": cmp["max_ref"], "roofline_frac": round(glups64 * 24 / HBM_PEAK_GBS, 4),
                       "t_block": s64["t_block"], "tile": [s64["tile_x"], s64["tile_y"]], "zchunk": s64["zchunk"]}
        st64.close()
    st32.close()
    if not args.no_extras:
        print("bench: extras", file=sys.stderr, flush=True)
        out["extras"] = {
            "jacobi7_f64_128": extra_first_order(1, 128, 100, coef=[0.4, 0.1], fp32=False, graph=True),
            "var7_nosym_512": extra_first_order(5, 512, 500, bf16=True),
            "batch16_wave25_bf16": extra_batch(500),
            "adjoint_wave25_400": extra_adjoint(),
        }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_nt)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
