"""Pins of the oracle's GR_FAST restatement (gro_sweep_fast_f32), on the CPU.

GR_FAST is the library's FP32 arithmetic for kernel 0 (radius sums with radii 1-2 as
differences from the centre, fmaf scaling, increment-form update); girih has no such
form, so no golden vector covers it.  The restatement is pinned by:
* the FP64 reference order (k0_wave25_f64, itself pinned by girih's golden vectors):
  one sweep and a 20-step run of a smooth field agree to a few FP32 ulps, while a
  dropped radius, a flipped sign or a transposed neighbour offset moves the result by
  orders of magnitude more;
* the exact Laplacian of a quadratic field (the 8th-order weights are exact to degree
  8): sum_r w_r * 2 r^2 = 2 per axis, so one sweep adds 6 * roc2 to a field at rest;
* the source injection and step pairing of gro_run_f32 (nb_naive_ts.c:190,196).
"""
import numpy as np
import pytest

import oracle as orc

W = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
COEF = np.array([3 * W[0]] + W[1:5])


def _smooth(nx, ny, nz, seed=3):
    shp = orc.shape_of(0, nx, ny, nz)
    kk, jj, ii = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shp], indexing="ij")
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.35, 0.65, 3) * np.array(shp[::-1])
    g = np.exp(-((ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2) / 40.0)
    g[:4], g[-4:], g[:, :4], g[:, -4:], g[:, :, :4], g[:, :, -4:] = 0, 0, 0, 0, 0, 0
    return g.astype(np.float32)


def _fp64_steps(g32, nt, roc2s, coef=COEF):
    u = g32.astype(np.float64)
    v = u.copy()
    c = np.zeros(11)
    c[:5] = coef
    orc.run(0, c, u, v, np.full(u.shape, roc2s), nt=nt)
    return u, v


def test_fast_one_sweep_close_to_fp64_reference_order():
    g = _smooth(40, 33, 29)
    u32, v32 = g.copy(), g.copy()
    orc.sweep_fast(COEF, u32, v32, roc2s=0.16)
    u64 = g.astype(np.float64)
    v64 = u64.copy()
    c = np.zeros(11)
    c[:5] = COEF
    orc.sweep(0, c, u64, v64, np.full(u64.shape, 0.16))
    err = np.abs(u32 - u64).max() / np.abs(u64).max()
    assert err < 2e-7, err


def test_fast_run_close_to_fp64_and_sensitive_to_planted_mistakes():
    nt, roc2s = 20, 0.16
    g = _smooth(36, 30, 28, seed=5)
    u32, v32 = g.copy(), g.copy()
    assert orc.run_fast(COEF, u32, v32, roc2s=roc2s, nt=nt) == nt
    _, v64 = _fp64_steps(g, nt, roc2s)
    rel = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    assert rel < 1e-6, rel
    # the tolerance separates plausible mistakes (planted in the FP64 reference here)
    for bad in (COEF * [1, 1, 1, 1, 0], COEF * [1, 1, -1, 1, 1], COEF * [1, 1, 1, 0, 1]):
        _, vb = _fp64_steps(g, nt, roc2s, bad)
        assert np.linalg.norm(v32 - vb) / np.linalg.norm(vb) > 100 * 1e-6


def test_fast_quadratic_laplacian_is_exact():
    nx, ny, nz = 12, 10, 9
    shp = orc.shape_of(0, nx, ny, nz)
    kk, jj, ii = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shp], indexing="ij")
    q = ((ii - 7) ** 2 + 2 * (jj - 5) ** 2 + 3 * (kk - 6) ** 2).astype(np.float32)  # lap = 2 + 4 + 6
    u = q.copy()
    orc.sweep_fast(COEF, u, q, roc2s=1.0)  # u = v + (v - u) + lap = v + lap for u = v
    inner = (slice(4, -4),) * 3
    lap = (u[inner].astype(np.float64) - q[inner])
    np.testing.assert_allclose(lap, 12.0, atol=2e-3)
    # per-cell roc2 field gives the same as the scalar
    u2 = q.copy()
    orc.sweep_fast(COEF, u2, q, roc2=np.ones_like(q))
    np.testing.assert_array_equal(u2, u)
    # an axis or sign mistake shows: the y-axis weight 2 alone would be off by 2 * 2 = 4
    q2 = ((jj - 5) ** 2).astype(np.float32)
    u3 = q2.copy()
    orc.sweep_fast(COEF, u3, q2, roc2s=1.0)
    np.testing.assert_allclose(u3[inner].astype(np.float64) - q2[inner], 2.0, atol=1e-3)


def test_fast_run_source_and_pairing():
    """run_fast = alternating sweeps with the source added after each (gro_run_f32's loop)."""
    nx, ny, nz = 20, 14, 12
    shp = orc.shape_of(0, nx, ny, nz)
    rng = np.random.default_rng(11)
    roc2 = rng.uniform(0.05, 0.2, shp).astype(np.float32)
    src = (np.sin(np.arange(6) * 0.9) + 0.2).astype(np.float32)
    pt = (4 + 7, 4 + 5, 4 + 6)
    u, v = np.zeros(shp, np.float32), np.zeros(shp, np.float32)
    assert orc.run_fast(COEF, u, v, roc2=roc2, nt=5, src_idx=pt, src=src) == 6
    a, b = np.zeros(shp, np.float32), np.zeros(shp, np.float32)
    idx = (pt[2], pt[1], pt[0])
    for s in range(6):
        if s % 2 == 0:
            orc.sweep_fast(COEF, a, b, roc2=roc2)
            a[idx] += src[s]
        else:
            orc.sweep_fast(COEF, b, a, roc2=roc2)
            b[idx] += src[s]
    np.testing.assert_array_equal(u, a)
    np.testing.assert_array_equal(v, b)
    assert np.abs(v).max() > 0


@pytest.mark.parametrize("shape", [(9, 9, 9), (17, 9, 13)])
def test_fast_halo_untouched(shape):
    nx, ny, nz = shape
    shp = orc.shape_of(0, nx, ny, nz)
    rng = np.random.default_rng(1)
    v = np.zeros(shp, np.float32)
    v[4:-4, 4:-4, 4:-4] = rng.uniform(0, 1, (nz, ny, nx))
    u = v.copy()
    u[:4] = 7.0  # halo values survive a sweep
    orc.sweep_fast(COEF, u, v, roc2s=0.1)
    assert np.all(u[:4] == 7.0)
