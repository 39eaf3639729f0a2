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


REF_BIN = {"dp": os.path.join(REPO, "oracle", "_ref", "mwd_kernel_dp"),
           "sp": os.path.join(REPO, "oracle", "_ref", "mwd_kernel_sp")}
RADIUS = {0: 4, 1: 1, 5: 1}


def host_threads():
    """Host cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS
    where the environment sets it (the CPU share the box grants one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    return (min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail), avail


def _girih(prec, kernel, shape, nt, ts, t_dim, threads, n_tests, timeout=900):
    """One run of girih's own mwd_kernel (built from its sources into oracle/_ref):
    returns (GLUP/s of the median test, nt as girih ran it, wall seconds)."""
    nx, ny, nz = shape
    cmd = [REF_BIN[prec], "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--nt", str(nt), "--target-ts", str(ts),
           "--target-kernel", str(kernel), "--thread-group-size", "1", "--thx", "1", "--thy", "1", "--thz", "1",
           "--num-wavefronts", "1", "--n-tests", str(n_tests), "--verbose", "1"]
    if ts == 2:
        cmd += ["--t-dim", str(t_dim)]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true", OMP_PLACES="cores",
               OMP_WAIT_POLICY="active")
    t0 = time.perf_counter()
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=timeout)
    wall = time.perf_counter() - t0
    tests = [float(x) for x in re.findall(r"Rank 0 TEST#\d+ time:\s*([0-9.eE+-]+)", r.stdout)]
    m = re.search(r"Number of time steps:\s*(\d+)", r.stdout)
    if r.returncode != 0 or len(tests) != n_tests or not m:
        raise RuntimeError("girih run failed (rc %d): %s | %s" % (r.returncode, " ".join(cmd[1:]),
                                                                   (r.stderr or r.stdout)[-300:]))
    nt_run = int(m.group(1))
    t_med = sorted(tests)[len(tests) // 2]  # performance.c:104-106
    return nx * ny * nz * nt_run / t_med / 1e9, nt_run, wall


def feasible_t_dims(kernel, ny, nz, threads, limit=4):
    """Odd diamond depths girih's stepper accepts for this grid (diamond width
    2R(t_dim+1) divides ny, nz >= 2 t_dim R + 1; diamond_utils.c:1060-1098) that
    leave at least one diamond per thread, smallest first."""
    R = RADIUS[kernel]
    out = []
    for t in range(1, 256, 2):
        dw = 2 * R * (t + 1)
        if ny % dw == 0 and nz >= 2 * t * R + 1 and ny // dw >= threads:
            out.append(t)
    return out[:limit] or [1]


def cpu_baseline(n, nt_c3, full=False):
    """girih's own OpenMP steppers on this box's host cores, following its
    performance_test protocol (median of 3 tests, performance.c:29-127): the 1WD
    diamond stepper (one thread per diamond, ts 2) with t_dim tuned over the feasible
    odd values, and spatial blocking (ts 0).  Configurations: the 25-point wave at
    n^3 in DP and SP (nt reduced to a stated sample; GLUP/s is steady state), the
    7-point Jacobi 128^3 x 100 (DP, in full) and the 7-point var-coef 512^3 (DP, nt
    reduced unless full).  Fails loudly (value null, error text) when girih is not
    built; there is no fallback."""
    threads, avail = host_threads()
    base = {"unit": "GLUP/s", "cores": threads, "kind": "reference", "nproc": os.cpu_count(),
            "affinity_cpus": avail, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    missing = [p for p in REF_BIN.values() if not os.path.exists(p)]
    if missing:
        msg = "girih's mwd_kernel is not built (make -C oracle ref needs its sources): %s" % missing
        print("bench.py: cpu_baseline: " + msg, file=sys.stderr, flush=True)
        return dict(base, value=None, error=msg, sample="not run")
    configs = {}
    try:
        def tuned(prec, kernel, shape, nt, tune_shape, tune_nt):
            cands = feasible_t_dims(kernel, shape[1], min(shape[2], tune_shape[2]), threads)
            trials, errors = {}, []
            for t in cands:  # tuning: one test each on a reduced z extent (same diamonds and wavefronts)
                try:
                    trials[t] = _girih(prec, kernel, tune_shape, tune_nt, 2, t, threads, 1)[0]
                except RuntimeError as e:  # a depth girih rejects for this grid is skipped
                    errors.append(str(e))
            if not trials:
                raise RuntimeError("no feasible diamond depth: %s" % errors)
            best = max(trials, key=trials.get)
            gl, nt_run, wall = _girih(prec, kernel, shape, nt, 2, best, threads, 3)
            return {"glups": round(gl, 3), "t_dim": best, "tuning_glups": {str(k): round(v, 3) for k, v in trials.items()},
                    "grid": list(shape), "nt": nt_run, "n_tests": 3, "wall_s": round(wall, 1)}

        print("bench: cpu baseline (girih 1WD / SB, %d threads)" % threads, file=sys.stderr, flush=True)
        c3 = (n, n, n)
        c3_tune = (n, n, min(n, 128))
        configs["wave25_sp_1wd"] = tuned("sp", 0, c3, nt_c3, c3_tune, 6)
        configs["wave25_dp_1wd"] = tuned("dp", 0, c3, nt_c3, c3_tune, 6)
        sb_shape = (n, n, max(8, n // 8))  # spatial blocking: 0.26 GLUP/s at 1024^3 on 16 cores
        gl, nt_run, wall = _girih("dp", 0, sb_shape, 2, 0, 1, threads, 3)
        configs["wave25_dp_sb"] = {"glups": round(gl, 3), "grid": list(sb_shape), "nt": nt_run, "n_tests": 3,
                                   "wall_s": round(wall, 1)}
        configs["jacobi7_dp_128_1wd"] = tuned("dp", 1, (128, 128, 128), 100, (128, 128, 128), 20)
        c2_nt = 500 if full else 50
        configs["var7_dp_512_1wd"] = tuned("dp", 5, (512, 512, 512), c2_nt, (512, 512, 64), 10)
    except (OSError, subprocess.SubprocessError, RuntimeError) as e:
        msg = "girih baseline failed: %s" % e
        print("bench.py: cpu_baseline: " + msg, file=sys.stderr, flush=True)
        return dict(base, value=None, error=msg, configs=configs, sample="failed")
    w = configs["wave25_sp_1wd"]
    return dict(base, value=w["glups"], configs=configs,
                sample="girih mwd_kernel built from its sources (SP, 1WD diamond stepper, t_dim %d tuned over %s), "
                       "25-pt wave %d^3 x nt %d, median of 3 tests on %d threads; DP, spatial blocking, the 7-pt "
                       "Jacobi 128^3 x 100 and the 7-pt var-coef 512^3 x %d in `configs`"
                       % (w["t_dim"], sorted(int(k) for k in w["tuning_glups"]), n, w["nt"], threads, c2_nt))


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
    if lowp:
        # The contracting stencil takes the field below FP32's range within the run (its
        # FP64 maximum is about 1e-61 after 500 steps), so the error is followed every 50
        # steps; the headline error is the last one taken while the FP64 field's maximum
        # is at least 1e-30, eight decades above FP32's smallest normal number.
        series = {k: [] for k in lowp}
        for st in keep.values():
            st.init_fields(1, 12345)
        done = 0
        while done < napply:
            m = min(50, napply - done)
            for st in keep.values():
                st.step(m)
            done += m
            f = (done - 1) % 2  # the field the last application wrote
            for k in lowp:
                c = keep[k].compare(f, keep["f64"], f)
                series[k].append((done, c["rel_l2"], c["max_ref"]))
        for k in lowp:
            live = [x for x in series[k] if x[2] >= 1e-30] or series[k][:1]
            out[k]["rel_l2_vs_f64"] = live[-1][1]
            out[k]["rel_l2_vs_f64_at_steps"] = live[-1][0]
            out[k]["f64_max_abs_at_those_steps"] = live[-1][2]
            out[k]["f64_max_abs_after_run"] = series[k][-1][2]
            out[k]["rel_l2_vs_f64_series"] = [[d, float("%.3g" % e), float("%.3g" % mr)] for d, e, mr in series[k]]
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
            "time_block": s["time_block"], "block_variant": s["block_variant"], "launches": s["last_launches"],
            "graph_first_call_ms": round(t_first, 2),
            # algorithmic bytes per update: single steps read u^(n-1), u^n and roc2 (bf16) and write
            # u^(n+1): 14 B; a two-step block reads 4 + 4 + 2 and writes 4 + 4 per two updates: 9 B
            "roofline_frac": round(gl * (9 if s["time_block"] == 2 else 14) / HBM_PEAK_GBS, 4),
            "bytes_per_lup": 9 if s["time_block"] == 2 else 14,
            "rel_l2_vs_f64_max": max(errs), "rel_l2_vs_f64": [float("%.3g" % e) for e in errs]}


def extra_adjoint(n=400, nt=2000, ckpt=50, nfd=20, seed=7):
    """Adjoint gradient of the 25-pt variable-velocity stepper: n^3, nt steps,
    h = 10 m, 8-layer velocity (+-5% blobs), 15 Hz Ricker at 50 m depth, receiver
    plane at z = 8 cells every step, misfit 0.5 ||d_pred - d_obs||^2 with d_obs from
    a +3% Gaussian-blob velocity perturbation; FP32 shot with checkpoints every
    `ckpt` steps through torch.autograd, checked against the FP64 GPU adjoint
    (bit-identical to the FP64 CPU adjoint at test sizes) and against central
    finite differences of the FP64 forward model on `nfd` cells the wave reached;
    the same for the mixed-precision shot (FP64 forward, FP32 backward)."""
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
    mixed = adjoint.MixedShot(n, n, n, nt, dth, src_pt, src, rec_k, ckpt=ckpt)
    s64, s32 = mixed.s64, mixed.s32  # its FP64 and FP32 shots serve the single-precision legs too
    try:
        d_obs = s64.forward(torch.tensor(m_obs, device="cuda"))
        m64 = torch.tensor(m0, device="cuda")
        tr64 = s64.forward(m64)
        g64 = s64.vjp(m64, tr64 - d_obs)
        # FP32: forward + backward through the autograd Function, timed after one untimed
        # pass (the first backward pass allocates its segment buffers)
        m32 = torch.tensor(m0, dtype=torch.float32, device="cuda", requires_grad=True)
        (0.5 * ((adjoint.shot_traces(m32, s32).double() - d_obs) ** 2).sum()).backward()
        m32.grad = None
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
        # mixed precision: FP64 forward at m0 (traces, residual), FP32 backward from its checkpoints
        mm = torch.tensor(m0, device="cuda", requires_grad=True)
        _sync()
        t3 = time.perf_counter()
        trm = adjoint.shot_traces(mm, mixed)
        _sync()
        t4 = time.perf_counter()
        lossm = 0.5 * ((trm - d_obs) ** 2).sum()
        lossm.backward()
        _sync()
        t5 = time.perf_counter()
        gmix = mm.grad
        rel_mix = float(torch.linalg.norm(gmix - g64) / torch.linalg.norm(g64))
        del trm
        J0 = float(0.5 * ((tr64 - d_obs) ** 2).sum())
        # the FP32 gradient error follows from the FP32 traces' error relative to the residual
        tr_err = float(torch.linalg.norm(tr32.detach().double() - tr64) / torch.linalg.norm(tr64))
        res_rel = float(torch.linalg.norm(tr64 - d_obs) / torch.linalg.norm(tr64))
        del tr32, tr64
        g = g64.cpu().numpy()
        g32h = g32.cpu().numpy()
        gmh = gmix.cpu().numpy()
        gmax = np.abs(g).max()
        live = np.argwhere(np.abs(g) > 1e-2 * gmax)
        rng = np.random.default_rng(seed)
        errs32, errs64, errs32n, errsm = [], [], [], []

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
            errsm.append(abs(gmh[k, j, i] - fd) / abs(fd))
    finally:
        mixed.close()
    fwd_glups = n ** 3 * nt / (t1 - t0) / 1e9
    return {"grid": n, "steps": nt, "checkpoint_every": ckpt, "forward_s": round(t1 - t0, 3),
            "backward_s": round(t2 - t1, 3), "forward_glups": round(fwd_glups, 2),
            "backward_over_forward": round((t2 - t1) / (t1 - t0), 3), "misfit": J0,
            "rel_l2_grad_fp32_vs_fp64": rel, "fd_cells": len(errs32),
            "fd_max_rel_err_fp32_adjoint": max(errs32), "fd_max_rel_err_fp64_adjoint": max(errs64),
            "fd_max_err_fp32_adjoint_over_max_grad": max(errs32n),
            "rel_l2_traces_fp32_vs_fp64": tr_err, "residual_over_traces": res_rel,
            "mixed": {"forward_s": round(t4 - t3, 3), "backward_s": round(t5 - t4, 3),
                      "fd_max_rel_err": max(errsm), "rel_l2_grad_vs_fp64": rel_mix,
                      "note": "FP64 m and forward run (traces, residual); FP32 recomputation and adjoint "
                              "from the FP64 checkpoints rounded to FP32; gradient returned as FP64"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-nt", type=int, default=20, help="time steps of the CPU baseline's 25-point wave sample")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline runs the 7-point var-coef 512^3 for all 500 steps")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps of the FP32-vs-FP64 error check (run apart from the timed steps when they differ)")
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
        E = args.error_steps
        if E != K:  # the error target is stated after 1000 steps: run them apart from the timed steps
            print("bench: FP32 vs FP64 error over %d steps" % E, file=sys.stderr, flush=True)
            ecfg = physics.wave_1024_config(nx=n, ny=n, nz=n, nt=E)
            for st in (st32, st64):
                st.init_fields(2)
                st.set_source(*ecfg["src_point"], ecfg["src"][:E + 2])
                st.step(E)
            field = (E - 1) % 2
            ecmp = st32.compare(field, st64, field)
        else:
            ecmp = cmp
        out["fp64"] = {"glups": round(glups64, 3), "ms_per_step": round(t64 * 1e3 / K, 4),
                       "fp32_over_fp64": round(glups32 / glups64, 3),
                       "rel_l2_fp32_vs_fp64": cmp["rel_l2"], "max_abs_err": cmp["max_diff"],
                       "max_abs_ref": cmp["max_ref"], "error_steps": E,
                       "rel_l2_fp32_vs_fp64_at_error_steps": ecmp["rel_l2"],
                       "max_abs_err_at_error_steps": ecmp["max_diff"], "max_abs_ref_at_error_steps": ecmp["max_ref"],
                       "targets_met": bool(ecmp["rel_l2"] <= 1e-5 and glups32 / glups64 >= 1.8),
                       "roofline_frac": round(glups64 * 24 / HBM_PEAK_GBS, 4),
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
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_nt, full=args.cpu_full)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
