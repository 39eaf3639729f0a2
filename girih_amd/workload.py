"""Run one stepper workload with a pinned launch configuration (for profilers).

    python -m girih_amd.workload --kernel 5 --n 512 --dtype f32 --config 0,256,0,6 --steps 12
    python -m girih_amd.workload --kernel 0 --n 1024 --dtype f64 --config 13,0,0,1 --steps 10

The fields start from the seeded uniform state (--init random: interior U(0,1);
coefficient arrays centre U[0.3,0.5), neighbours U[0,0.1); kernel 0: the 1024^3
wave set-up, v = 3000 m/s, h = 10 m, CFL 0.4, scalar roc2), from zero fields
(--init zero) or from girih's closed-form state (--init girih).  --config is
VARIANT,ZCHUNK,XCD_ORDER,T_BLOCK as in gr_set_config / gr_set_xcd_order; with no
--config the autotuner picks one.  One line of JSON: the configuration, the
applications per launch and the GLUP/s of the timed steps (HIP events).
"""
import argparse
import json

from . import gpu, physics


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", type=int, required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--bf16", action="store_true", help="coefficient arrays / roc2 stored as bf16")
    ap.add_argument("--config", default=None)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--init", choices=["random", "zero", "girih"], default="random")
    a = ap.parse_args()
    dtype = gpu.GR_F32 if a.dtype == "f32" else gpu.GR_F64
    flags = gpu.GR_COEF_BF16 if a.bf16 else 0
    coef = None
    if a.kernel == 0:
        flags |= gpu.GR_ROC2_SCALAR | (gpu.GR_FAST if dtype == gpu.GR_F32 else 0)
        coef = physics.fd8_coef()
    elif a.kernel == 1:
        coef = [0.4, 0.1]
    st = gpu.Stepper(a.kernel, a.n, a.n, a.n, dtype=dtype, flags=flags, roc2=0.16, coef=coef)
    if gpu.NUM_COEF_ARRAYS[a.kernel]:
        st.init_coef_uniform(5, 0.3, 0.5, 0.0, 0.1)
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
    st.step(napply)
    s = st.stats()
    print(json.dumps({"kernel": a.kernel, "n": a.n, "dtype": a.dtype, "bf16": a.bf16, "init": a.init,
                      "config": "%d,%d,%d,%d" % (s["tile_variant"], s["zchunk"], s["xcd_order"], s["t_block"]),
                      "applications": napply, "launches": s["last_launches"],
                      "glups": a.n ** 3 * napply / s["last_ms"] / 1e6}))
    st.close()


if __name__ == "__main__":
    main()
