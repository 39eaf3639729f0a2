"""The FP64 CPU shot/adjoint (oracle/adjoint.c) pinned by mathematics.

girih has no adjoint, so the CPU reference for the gradient is pinned by what the
mathematics fixes: the gradient of J(m) = 0.5 ||trace(m) - d_obs||^2 must equal
central finite differences of J itself (a dropped term, a sign, an off-by-one time
index or a missing source correction all break it), the forward traces must equal
the oracle's own time loop (which is bit-identical to girih's reference), and the
adjoint must be linear in the trace cotangent.
"""
import numpy as np
import pytest

import oracle as orc
from girih_amd import physics

NX, NY, NZ, NT = 12, 10, 14, 36
DT_OVER_H = 0.4 / 4725.0
COEF = physics.fd8_coef()


def _model(seed):
    rng = np.random.default_rng(seed)
    shp = orc.shape_of(0, NX, NY, NZ)
    z = np.arange(shp[0])[:, None, None]
    v = 1500.0 + 3000.0 * (z / shp[0]) + 200.0 * rng.standard_normal(shp) * 0.1
    return (v * DT_OVER_H) ** 2  # roc2


def _setup(seed=1):
    roc2 = _model(seed)
    src = physics.ricker(90.0, 0.4 * 10 / 4725.0, NT, t0=0.006)
    src_idx = (4 + NX // 2, 4 + NY // 2, 4 + 3)
    rec_k = 4 + 5
    return roc2, src, src_idx, rec_k


def _J(roc2, src, src_idx, rec_k, d_obs):
    tr, _ = orc.shot(COEF, roc2, NT, src_idx, src, rec_k)
    return 0.5 * np.sum((tr - d_obs) ** 2)


def test_shot_traces_equal_oracle_time_loop():
    roc2, src, src_idx, rec_k = _setup()
    tr, _ = orc.shot(COEF, roc2, NT, src_idx, src, rec_k)
    shp = roc2.shape
    u = np.zeros(shp)
    v = np.zeros(shp)
    c = np.zeros(11)
    c[:5] = COEF
    for a in range(NT):  # x_{a+2} overwrites x_a: the oracle's u/v pairing
        if a % 2 == 0:
            orc.sweep(0, c, u, v, roc2)
            u[src_idx[2], src_idx[1], src_idx[0]] += src[a]
            x = u
        else:
            orc.sweep(0, c, v, u, roc2)
            v[src_idx[2], src_idx[1], src_idx[0]] += src[a]
            x = v
        np.testing.assert_array_equal(tr[a], x[rec_k, 4:-4, 4:-4])
    assert np.abs(tr).max() > 0


def test_adjoint_gradient_matches_central_differences():
    roc2, src, src_idx, rec_k = _setup()
    d_obs = orc.shot(COEF, _model(7), NT, src_idx, src, rec_k)[0]  # data of another model
    tr, _ = orc.shot(COEF, roc2, NT, src_idx, src, rec_k)
    _, grad = orc.shot(COEF, roc2, NT, src_idx, src, rec_k, res=tr - d_obs, dt_over_h=DT_OVER_H)
    m = DT_OVER_H ** 2 / roc2
    rng = np.random.default_rng(3)
    cells = [(rng.integers(0, NZ), rng.integers(0, NY), rng.integers(0, NX)) for _ in range(8)]
    cells += [(src_idx[2] - 4, src_idx[1] - 4, src_idx[0] - 4),  # the source cell
              (rec_k - 4, 3, 5), (rec_k - 4 + 1, NY // 2, NX // 2)]  # on / next to the receiver plane
    for (k, j, i) in cells:
        h = 1e-5 * m[k + 4, j + 4, i + 4]
        rp = roc2.copy()
        rm = roc2.copy()
        rp[k + 4, j + 4, i + 4] = DT_OVER_H ** 2 / (m[k + 4, j + 4, i + 4] + h)
        rm[k + 4, j + 4, i + 4] = DT_OVER_H ** 2 / (m[k + 4, j + 4, i + 4] - h)
        fd = (_J(rp, src, src_idx, rec_k, d_obs) - _J(rm, src, src_idx, rec_k, d_obs)) / (2 * h)
        g = grad[k, j, i]
        scale = np.abs(grad).max()
        print("cell", (k, j, i), "adjoint", g, "fd", fd, flush=True)
        assert abs(g - fd) <= 1e-6 * scale + 1e-5 * abs(fd)


def test_adjoint_is_linear_in_the_cotangent():
    roc2, src, src_idx, rec_k = _setup()
    rng = np.random.default_rng(5)
    r1 = rng.standard_normal((NT, NY, NX))
    r2 = rng.standard_normal((NT, NY, NX))
    g1 = orc.shot(COEF, roc2, NT, src_idx, src, rec_k, res=r1, dt_over_h=DT_OVER_H)[1]
    g2 = orc.shot(COEF, roc2, NT, src_idx, src, rec_k, res=r2, dt_over_h=DT_OVER_H)[1]
    g12 = orc.shot(COEF, roc2, NT, src_idx, src, rec_k, res=2.0 * r1 - r2, dt_over_h=DT_OVER_H)[1]
    np.testing.assert_allclose(g12, 2.0 * g1 - g2, rtol=1e-9, atol=1e-12 * np.abs(g1).max())


def test_shot_rejects_bad_receiver_plane():
    roc2, src, src_idx, _ = _setup()
    with pytest.raises(ValueError):
        orc.shot(COEF, roc2, NT, src_idx, src, 2)
