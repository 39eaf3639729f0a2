"""Physical set-ups for the 25-point wave stepper (kernel 0).

girih's own runs use synthetic coefficients (utils.c:116-119: 1-D 8th-order
weights / 10 with the 1-D centre weight) and a linear ramp for roc2
(utils.c:638-644); its point source is disabled (utils.c:336).  The physical
workloads below map onto the same kernel:

    u^(n+1) = 2u^n - u^(n-1) + roc2 * Lap8(u^n),   roc2 = (v dt / h)^2

with girih's coefficient slots coef[0] = 3*w0 (3-D centre), coef[r] = w_r, and a
Ricker wavelet injected after every step at the source cell (nb_naive_ts.c:190,196).
"""
import math

import numpy as np

# 8th-order central second-derivative weights (offsets 0..4)
FD8 = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)


def fd8_coef():
    """girih coef[0..4] for the isotropic 3-D 8th-order Laplacian (h = 1)."""
    return [3.0 * FD8[0]] + list(FD8[1:])


def cfl_dt(vmax, h, cfl=0.4):
    return cfl * h / vmax


def roc2_const(v, dt, h):
    return (v * dt / h) ** 2


def ricker(f0, dt, n, t0=None):
    """Ricker wavelet samples r(t_s), t_s = s*dt, s < n; peak at t0 (default 1/f0)."""
    t0 = 1.0 / f0 if t0 is None else t0
    t = np.arange(n) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def wave_1024_config(nx=1024, ny=1024, nz=1024, nt=1000, v=3000.0, h=10.0, cfl=0.4, f0=15.0):
    """The 25-point constant-velocity wave workload: v = 3000 m/s, h = 10 m, dt at
    CFL 0.4, 15 Hz Ricker at the grid centre; returns a dict of kernel inputs."""
    dt = cfl_dt(v, h, cfl)
    r2 = roc2_const(v, dt, h)
    R = 4
    # source amplitude scaled like a pressure injection (v dt)^2 / h^3 * h^3 -> roc2
    src = r2 * ricker(f0, dt, nt + 2)
    return {
        "nx": nx, "ny": ny, "nz": nz, "nt": nt, "dt": dt, "h": h, "v": v, "roc2": r2, "coef": fd8_coef(),
        "src_point": (R + nx // 2, R + ny // 2, R + nz // 2), "src": src, "f0": f0,
    }
