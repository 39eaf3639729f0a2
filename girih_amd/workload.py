"""Run one stepper workload with a pinned launch configuration (for profilers).

    python -m girih_amd.workload --kernel 5 --n 512 --dtype f32 --config 0,256,0,6 --steps 12
    python -m girih_amd.workload --kernel 0 --n 1024 --dtype f64 --config 13,0,0,1 --steps 10
    python -m girih_amd.workload --batch --config 1,64 --steps 20

The fields start from the seeded uniform state (--init random: interior U(0,1);
coefficient arrays centre U[0.3,0.5), neighbours U[0,0.1); kernel 0: the 1024^3
wave set-up, v = 3000 m/s, h = 10 m, CFL 0.4, scalar roc2; with --roc2-field or
--bf16, the batch's 8-layer velocity model as a per-cell roc2), from zero fields
(--init zero) or from girih's closed-form state (--init girih).  --config is
VARIANT,ZCHUNK,XCD_ORDER,T_BLOCK as in gr_set_config / gr_set_xcd_order; with no
--config the autotuner picks one.  One line of JSON: the configuration, the
applications per launch and the GLUP/s of the timed steps (HIP events), which follow
one untimed pass (--warmup 0 drops it).

--batch runs the 16-problem wave batch instead (kernel 0, FP32 GR_FAST, bf16 roc2
from 8-layer velocity models, dims uniform in 192..512 step 16 from seed 2024, a
Ricker source at each centre), with --config VARIANT,ZCHUNK[,TBLOCK[,BLOCK_VARIANT]]
of gr_batch_set_config and gr_batch_set_time_block, or the batch autotuner; one
persistent launch per step (TBLOCK 1) or per two-step block (2), replayed from a HIP
graph (--mode 1, the default), launched one by one (--mode 0), or all in one
cooperative launch (--mode 2).

--shot N,NT,CKPT runs one seismic shot instead (girih_amd.adjoint; 8-layer model,
15 Hz Ricker at 50 m depth, receivers at z = 8 cells): the forward run and the
vector-Jacobian product of a seeded cotangent, twice, timing the second
(--shot-prec f32, f64 or mixed).
"""
import argparse
import json

from . import gpu, physics


def batch_main(a):
    import numpy as np
    rng = np.random.default_rng(2024)
    sizes = [tuple(int(x) for x in rng.choice(np.arange(192, 513, 16), 3)) for _ in range(16)]
    h = 10.0
    dt = physics.cfl_dt(4500.0 * 1.05, h, 0.4)
    src = physics.roc2_const(3000.0, dt, h) * physics.ricker(15.0, dt, a.steps + 16)
    cells = sum(x * y * z for x, y, z in sizes)
    with gpu.Batch(sizes, flags=gpu.GR_FAST | gpu.GR_COEF_BF16, coef=physics.fd8_coef()) as b:
        traces = []
        for i, p in enumerate(b.problems):
            p.init_fields(2)
            p.init_roc2_model(gpu.vmodel(2024 + i, dt / h))
            nx, ny, nz = sizes[i]
            p.set_source(4 + nx // 2, 4 + ny // 2, 4 + nz // 2, src)
            if a.receivers:  # a receiver plane 8 cells below the top, every application
                import torch
                traces.append(torch.empty((a.steps + 2, ny, nx), dtype=torch.float32, device="cuda"))
                b.set_receivers(i, 4 + 8, traces[-1], a.steps + 2)
        if a.config:
            cfg = [int(x) for x in a.config.split(",")] + [1, -1]
            b.set_config(cfg[0], cfg[1])
            b.set_time_block(cfg[2], cfg[3])
        else:
            b.autotune(4, mode=a.mode)
        b.step(2, mode=a.mode)
        b.step(a.steps, mode=a.mode)
        s = b.stats()
        print(json.dumps({"batch": len(sizes), "cells": cells,
                          "config": "%d,%d,%d,%d" % (s["variant"], s["zchunk"], s["time_block"], s["block_variant"]),
                          "tile": [s["tile_x"], s["tile_y"]], "steps": a.steps, "launches": s["last_launches"],
                          "mode": s["last_mode"], "grid": s["grid"], "receivers": a.receivers,
                          "glups": cells * a.steps / s["last_ms"] / 1e6}))


def shot_main(a):
    """One shot (girih_amd.adjoint): forward traces, then the vector-Jacobian product of
    a seeded cotangent; FP32 (GR_FAST), FP64 or mixed precision."""
    import time

    import numpy as np
    import torch

    from . import adjoint
    n, nt, ckpt = (int(x) for x in a.shot.split(","))
    h = 10.0
    dt = physics.cfl_dt(4500.0 * 1.05, h, 0.4)
    src = physics.roc2_const(1500.0, dt, h) * physics.ricker(15.0, dt, nt)
    pt = (4 + n // 2, 4 + n // 2, 4 + 5)
    with gpu.Stepper(0, n, n, n, dtype=gpu.GR_F64, coef=physics.fd8_coef()) as st:
        st.init_roc2_model(gpu.vmodel(7, dt / h))
        m0 = np.ascontiguousarray((dt / h) ** 2 / st.download_roc2()[4:-4, 4:-4, 4:4 + n])
    if a.shot_prec == "mixed":
        shot = adjoint.MixedShot(n, n, n, nt, dt / h, pt, src, 4 + 8, ckpt=ckpt)
        m = torch.tensor(m0, device="cuda")
    else:
        dtype = gpu.GR_F32 if a.shot_prec == "f32" else gpu.GR_F64
        if dtype == gpu.GR_F32:
            src = src.astype(np.float32).astype(np.float64)
        shot = adjoint.Shot(n, n, n, nt, dt / h, pt, src, 4 + 8, ckpt=ckpt, dtype=dtype)
        m = torch.tensor(m0, dtype=shot.torch_dtype, device="cuda")
    res = None
    times = []
    for _ in range(2):  # the first round loads the code objects and allocates the backward buffers
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = shot.forward(m)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if res is None:
            g = torch.Generator(device="cuda").manual_seed(11)
            res = torch.randn(tr.shape, generator=g, device="cuda", dtype=tr.dtype)
        shot.vjp(m, res)
        torch.cuda.synchronize()
        times.append((t1 - t0, time.perf_counter() - t1))
    shot.close()
    print(json.dumps({"shot": [n, n, n], "steps": nt, "checkpoint_every": ckpt, "precision": a.shot_prec,
                      "forward_s": round(times[-1][0], 4), "backward_s": round(times[-1][1], 4)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", action="store_true", help="the 16-problem bf16 wave batch")
    ap.add_argument("--shot", default=None, help="N,NT,CKPT: an N^3 shot of NT steps, forward and adjoint")
    ap.add_argument("--receivers", action="store_true", help="--batch: record a receiver plane of every problem")
    ap.add_argument("--shot-prec", choices=["f32", "f64", "mixed"], default="f32")
    ap.add_argument("--kernel", type=int)
    ap.add_argument("--n", type=int)
    ap.add_argument("--shape", default=None, help="NX,NY,NZ (instead of --n for all three)")
    ap.add_argument("--strm", type=int, choices=[-1, 0, 1], default=-1,
                    help="cache policy (gr_set_streaming): -1 by working-set size, 0 normal, 1 non-temporal")
    ap.add_argument("--graph", action="store_true", help="replay the timed steps from a HIP graph")
    ap.add_argument("--mode", type=int, choices=[0, 1, 2], default=1,
                    help="batch launches (gr_batch_step): 0 per step, 1 graph replay, 2 one cooperative launch")
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--bf16", action="store_true",
                    help="coefficient arrays / roc2 stored as bf16 (kernel 0: a bf16 roc2 field)")
    ap.add_argument("--roc2-field", action="store_true", help="kernel 0: a per-cell roc2 field (layered model)")
    ap.add_argument("--config", default=None)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--init", choices=["random", "zero", "girih"], default="random")
    ap.add_argument("--warmup", type=int, choices=[0, 1], default=1,
                    help="one untimed pass first, so the timed steps exclude the first launch's code-object load")
    a = ap.parse_args()
    if a.batch:
        return batch_main(a)
    if a.shot:
        return shot_main(a)
    if a.kernel is None or (a.n is None and a.shape is None):
        ap.error("--kernel and --n (or --shape) are required (or --batch)")
    nx, ny, nz = (int(x) for x in a.shape.split(",")) if a.shape else (a.n, a.n, a.n)
    dtype = gpu.GR_F32 if a.dtype == "f32" else gpu.GR_F64
    flags = gpu.GR_COEF_BF16 if a.bf16 else 0
    coef = None
    roc2_field = a.kernel == 0 and (a.bf16 or a.roc2_field)
    if a.kernel == 0:
        flags |= (0 if roc2_field else gpu.GR_ROC2_SCALAR) | (gpu.GR_FAST if dtype == gpu.GR_F32 else 0)
        coef = physics.fd8_coef()
    elif a.kernel == 1:
        coef = [0.4, 0.1]
    st = gpu.Stepper(a.kernel, nx, ny, nz, dtype=dtype, flags=flags, roc2=0.16, coef=coef)
    st.set_streaming(a.strm)
    if gpu.NUM_COEF_ARRAYS[a.kernel]:
        st.init_coef_uniform(5, 0.3, 0.5, 0.0, 0.1)
    if roc2_field:  # the batch's 8-layer velocity model (h = 10 m, CFL 0.4 at 4725 m/s)
        st.init_roc2_model(gpu.vmodel(2024, physics.cfl_dt(4500.0 * 1.05, 10.0, 0.4) / 10.0))
    st.init_fields({"girih": 0, "random": 1, "zero": 2}[a.init], 12345)
    if a.config:
        var, zc, order, tb = (int(x) for x in a.config.split(","))
        st.set_config(var, zc, tb)
        st.set_xcd_order(order)
    else:
        st.autotune(4)
    s = st.stats()
    tb = s["t_block"]
    # whole fused passes (first-order kernels end a step count with one single sweep)
    napply = a.steps * tb + (0 if a.kernel == 0 or tb == 1 else 1)
    if a.warmup:
        st.step(tb if a.kernel == 0 or tb == 1 else tb + 1)  # loads the code objects
    if a.graph:
        napply += napply % 2  # an even count restores the parity, so the timed call replays the capture
        st.set_graph(True)
        st.step(napply)  # captures the sequence
    st.step(napply)
    s = st.stats()
    print(json.dumps({"kernel": a.kernel, "shape": [nx, ny, nz], "dtype": a.dtype, "bf16": a.bf16, "init": a.init,
                      "config": "%d,%d,%d,%d" % (s["tile_variant"], s["zchunk"], s["xcd_order"], s["t_block"]),
                      "strm": a.strm, "graph": a.graph, "applications": napply, "launches": s["last_launches"],
                      "ms": s["last_ms"], "glups": nx * ny * nz * napply / s["last_ms"] / 1e6}))
    st.close()


if __name__ == "__main__":
    main()
