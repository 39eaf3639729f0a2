"""Temporally blocked (fused) passes: several time steps per HBM pass.

The fused kernels keep girih's operand order, so FP64/FP32 results must stay
bit-identical to the CPU oracle (and to single sweeps) for every depth, ragged
size, step count and source -- including the final U1/U2 pair.

The runs without a source are also the regression test of a fixed fault: the
kernel-1 pass visits planes below 0 in its first z chunk, where the no-source
marker (plane -1) once matched and a null source pointer was read.
"""
import numpy as np
import pytest

import oracle as orc
from girih_amd import gpu

from test_gpu_parity import _bf16_round, _consts, _pad, _pad_coef, _random_problem

pytestmark = pytest.mark.gpu

SIZES = {1: [(70, 38, 20), (130, 9, 11), (64, 32, 8)], 4: [(72, 20, 17), (67, 33, 12)]}


@pytest.mark.parametrize("kernel,depth", [(1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (5, 2), (5, 4), (5, 6), (0, 2)])
@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("strm", [0, 1])
def test_fused_bitexact_vs_oracle(kernel, depth, prec, strm):
    """strm 1 forces the non-temporal instantiations of the kernel-0 and kernel-1 passes
    (the ones problems larger than the caches run; kernel 5's pass has one form)."""
    if strm and kernel == 5:
        pytest.skip("the kernel-5 pass has no non-temporal form")
    dt = np.float64 if prec == "dp" else np.float32
    dtype = gpu.GR_F64 if prec == "dp" else gpu.GR_F32
    for n, (nx, ny, nz) in enumerate(SIZES[orc.RADIUS[kernel]]):
        for napply in (2 * depth + 1, 3 * depth):
            print("fused", kernel, depth, prec, strm, (nx, ny, nz), napply, flush=True)
            u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 1000 + 10 * kernel + n)
            with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel)) as st:
                st.set_streaming(strm)
                st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
                st.set_config(t_block=depth)
                st.step(napply)
                s = st.stats()
                assert s["last_launches"] < napply, "fused passes were not used"
                U1, U2 = st.download(_pad(u).shape)
            for a in range(napply):
                if a % 2 == 0:
                    orc.sweep(kernel, coef, u, v, roc2)
                else:
                    orc.sweep(kernel, coef, v, u, roc2)
            nnx = u.shape[2]
            np.testing.assert_array_equal(U1[:, :, :nnx], u)
            np.testing.assert_array_equal(U2[:, :, :nnx], v)


@pytest.mark.parametrize("depth", [2, 4, 6])
@pytest.mark.parametrize("zchunk", [0, 3])
def test_fused_bf16_coefficients_bitexact_vs_oracle(depth, zchunk):
    """Kernel 5 with its seven coefficient arrays stored as bf16 (GR_COEF_BF16):
    fused passes widen each coefficient exactly to FP32 and keep girih's operand
    order, so they are bit-identical to the FP32 oracle run on the bf16-rounded
    coefficients, for whole-z and 3-plane chunks."""
    kernel = 5
    for n, (nx, ny, nz) in enumerate(SIZES[1][:2]):
        napply = 2 * depth + 1
        print("fused bf16", depth, zchunk, (nx, ny, nz), flush=True)
        u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, np.float32, 700 + n)
        coef = _bf16_round(coef)
        with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_COEF_BF16) as st:
            st.upload(_pad(u), _pad(v), None, _pad_coef(coef, kernel))
            st.set_config(zchunk=zchunk, t_block=depth)
            st.step(napply)
            assert st.stats()["last_launches"] < napply, "fused passes were not used"
            U1, U2 = st.download(_pad(u).shape)
        for a in range(napply):
            if a % 2 == 0:
                orc.sweep(kernel, coef, u, v, roc2)
            else:
                orc.sweep(kernel, coef, v, u, roc2)
        nnx = u.shape[2]
        np.testing.assert_array_equal(U1[:, :, :nnx], u)
        np.testing.assert_array_equal(U2[:, :, :nnx], v)


@pytest.mark.parametrize("zchunk", [0, 5])
def test_bf16_roc2_wave_two_step_passes(zchunk):
    """The two-step wave pass reads a bf16 roc2 field widened exactly to FP32: two
    passes for four steps, bit-identical to the oracle on the bf16 values."""
    kernel, (nx, ny, nz) = 0, SIZES[4][0]
    u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, np.float32, 733)
    roc2 = _bf16_round(roc2)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_COEF_BF16, coef=_consts(coef, kernel)) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2), None)
        st.set_config(zchunk=zchunk, t_block=2)
        st.step(4)
        assert st.stats()["last_launches"] == 2
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, roc2, nt=4)
    nnx = u.shape[2]
    np.testing.assert_array_equal(U1[:, :, :nnx], u)
    np.testing.assert_array_equal(U2[:, :, :nnx], v)


def test_scalar_roc2_with_bf16_flag_two_step_passes():
    """GR_COEF_BF16 with a scalar roc2 has no field to store as bf16: the two-step
    pass runs the scalar form and matches the oracle's GR_FAST restatement."""
    nx, ny, nz = 70, 21, 13
    rng = np.random.default_rng(92)
    shp = orc.shape_of(0, nx, ny, nz)
    u = np.zeros(shp, np.float32)
    v = np.zeros(shp, np.float32)
    u[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    v[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef = [3 * w[0]] + w[1:5]
    flags = gpu.GR_FAST | gpu.GR_ROC2_SCALAR | gpu.GR_COEF_BF16
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F32, flags=flags, roc2=0.12, coef=coef) as st:
        st.upload(_pad(u), _pad(v))
        st.set_config(t_block=2)
        st.step(4)
        assert st.stats()["last_launches"] == 2
        U1, U2 = st.download(_pad(u).shape)
    orc.run_fast(coef, u, v, roc2s=0.12, nt=4)
    np.testing.assert_array_equal(U1[:, :, :shp[2]], u)
    np.testing.assert_array_equal(U2[:, :, :shp[2]], v)


def test_bf16_roc2_fast_two_step_bitexact_vs_oracle():
    """GR_FAST | GR_COEF_BF16 (the batch's arithmetic and storage) in two-step
    passes, with a source: bit-identical to the oracle's GR_FAST restatement run on
    the bf16 roc2 values."""
    nx, ny, nz = 150, 37, 21
    rng = np.random.default_rng(91)
    shp = orc.shape_of(0, nx, ny, nz)
    u = np.zeros(shp, np.float32)
    v = np.zeros(shp, np.float32)
    u[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    v[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    roc2 = np.zeros(shp, np.float32)
    roc2[4:-4, 4:-4, 4:-4] = _bf16_round(rng.uniform(0.02, 0.16, (nz, ny, nx)).astype(np.float32))
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef = [3 * w[0]] + w[1:5]
    src = np.linspace(1.0, -1.0, 9)
    pt = (4 + 64, 4 + 16, 4 + 9)
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_COEF_BF16, coef=coef) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2), None)
        st.set_source(*pt, src)
        st.set_config(zchunk=6, t_block=2)
        st.step(10)
        assert st.stats()["last_launches"] == 5
        U1, U2 = st.download(_pad(u).shape)
    src10 = np.concatenate([src, [0.0]]).astype(np.float32)  # the library adds nothing past its 9 values
    orc.run_fast(coef, u, v, roc2=roc2, nt=10, src_idx=pt, src=src10)
    np.testing.assert_array_equal(U1[:, :, :shp[2]], u)
    np.testing.assert_array_equal(U2[:, :, :shp[2]], v)


@pytest.mark.parametrize("kernel,depth", [(1, 2), (1, 3), (1, 5), (1, 6), (5, 2), (5, 4), (0, 2)])
@pytest.mark.parametrize("zchunk", [1, 5, 8])
@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_fused_zchunks_bitexact_vs_oracle(kernel, depth, zchunk, prec):
    """Workgroups that own a z chunk recompute the chunk's z ghost planes; the
    result must not depend on the chunk size (including chunks of one plane)."""
    nx, ny, nz = SIZES[orc.RADIUS[kernel]][0]
    napply = 2 * depth + (0 if kernel == 0 else 1)
    print("fused zchunk", kernel, depth, zchunk, prec, flush=True)
    dt, dtype = (np.float64, gpu.GR_F64) if prec == "dp" else (np.float32, gpu.GR_F32)
    u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 500 + kernel)
    outs = []
    with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel)) as st:
        for order in (0, 1, 2):  # plain, XCD-contiguous and per-layer XCD-band workgroup order
            st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
            st.set_config(zchunk=zchunk, t_block=depth)
            st.set_xcd_order(order)
            st.step(napply)
            assert st.stats()["last_launches"] < napply
            outs.append(st.download(_pad(u).shape))
    for a in range(napply):
        if a % 2 == 0:
            orc.sweep(kernel, coef, u, v, roc2)
        else:
            orc.sweep(kernel, coef, v, u, roc2)
    nnx = u.shape[2]
    for order, (U1, U2) in enumerate(outs):
        np.testing.assert_array_equal(U1[:, :, :nnx], u, err_msg="order %d" % order)
        np.testing.assert_array_equal(U2[:, :, :nnx], v, err_msg="order %d" % order)


def test_fused_wave_with_source_fp64():
    nx, ny, nz = 66, 40, 30
    kernel = 0
    shp = orc.shape_of(kernel, nx, ny, nz)
    u = np.zeros(shp)
    v = np.zeros(shp)
    roc2 = np.full(shp, 0.16)
    coef = np.zeros(11)
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef[0] = 3 * w[0]
    coef[1:5] = w[1:5]
    src = np.cos(np.arange(16) * 0.9) + 0.2
    pt = (4 + 40, 4 + 20, 4 + 15)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=list(coef[:5])) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2))
        st.set_source(*pt, src[:13])
        st.set_config(t_block=2)
        st.step(16)  # source values past 13 add nothing
        U1, U2 = st.download(_pad(u).shape)
    s2 = np.zeros(16)
    s2[:13] = src[:13]
    orc.run(kernel, coef, u, v, roc2, nt=16, src_idx=pt, src=s2)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("roc2_kind", ["field", "scalar"])
def test_fused_wave_tiles_and_source_on_tile_edges(prec, roc2_kind):
    """Two-step kernel-0 passes over several xy tiles and z chunks, with the point
    source on the first column and row of a tile (a margin cell of the neighbouring
    tiles' first level): bit-identical to the oracle, with a roc2 field and with a
    scalar roc2 (which the margin cells must use as well)."""
    nx, ny, nz = 300, 50, 21
    kernel = 0
    dt, dtype = (np.float64, gpu.GR_F64) if prec == "dp" else (np.float32, gpu.GR_F32)
    u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 4242)
    kw = {}
    if roc2_kind == "scalar":
        roc2 = np.full_like(roc2, 0.13)
        kw = dict(flags=gpu.GR_ROC2_SCALAR, roc2=0.13)
    pt = (4 + (64 if prec == "dp" else 128), 4 + 16, 4 + 10)  # output tiles are 64 x 16 (FP64), 128 x 16 (FP32)
    src = np.cos(np.arange(10) * 0.7)
    print("fused wave tiles", prec, roc2_kind, flush=True)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel), **kw) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2) if roc2_kind == "field" else None)
        st.set_source(*pt, src)
        st.set_config(zchunk=5, t_block=2)
        st.step(10)
        assert st.stats()["last_launches"] == 5
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, roc2, nt=10, src_idx=pt, src=src)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


@pytest.mark.parametrize("strm", [0, 1])
@pytest.mark.parametrize("zchunk", [0, 3, 7])
def test_fused_fast_bitexact_vs_oracle(strm, zchunk):
    """GR_FAST two-step wave passes with a scalar roc2 (the benchmark's arithmetic and
    fused form), over several tiles and z chunks, with a point source on a tile edge:
    bit-identical to the oracle's restatement of the GR_FAST arithmetic."""
    nx, ny, nz = 200, 37, 23
    kernel = 0
    rng = np.random.default_rng(90 + zchunk)
    shp = orc.shape_of(kernel, nx, ny, nz)
    u = np.zeros(shp, np.float32)
    v = np.zeros(shp, np.float32)
    u[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    v[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef = [3 * w[0]] + w[1:5]
    src = np.linspace(1.0, -1.0, 10)
    pt = (4 + 64, 4 + 16, 4 + 9)
    print("fused fast", strm, zchunk, flush=True)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_ROC2_SCALAR, roc2=0.12,
                     coef=coef) as st:
        st.set_streaming(strm)
        st.upload(_pad(u), _pad(v))
        st.set_source(*pt, src)
        st.set_config(zchunk=zchunk, t_block=2)
        st.step(10)
        assert st.stats()["last_launches"] == 5
        U1, U2 = st.download(_pad(u).shape)
    orc.run_fast(coef, u, v, roc2s=0.12, nt=10, src_idx=pt, src=src.astype(np.float32))
    np.testing.assert_array_equal(U1[:, :, :shp[2]], u)
    np.testing.assert_array_equal(U2[:, :, :shp[2]], v)


def test_fused_fast_fp32_multi_tile_equals_single_sweeps():
    """GR_FAST two-step passes with a scalar roc2 over several tiles, z chunks and a
    source on a tile edge equal GR_FAST single sweeps bit for bit."""
    nx, ny, nz = 260, 40, 30
    kernel = 0
    u, v, _, coef = _random_problem(kernel, nx, ny, nz, np.float32, 78)
    outs = []
    for depth in (1, 2):
        with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_ROC2_SCALAR, roc2=0.12,
                         coef=list(coef[:5])) as st:
            st.upload(_pad(u), _pad(v))
            st.set_source(4 + 127, 4 + 15, 4 + 7, np.linspace(1.0, -1.0, 12))
            st.set_config(zchunk=7, t_block=depth)
            st.step(12)
            outs.append(st.download(_pad(u).shape))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("kernel,depth", [(1, 2), (1, 4), (1, 6), (5, 4), (0, 2)])
def test_fused_tiny_grids_bitexact(kernel, depth):
    """Fused passes on grids smaller than one tile and than the pipeline depth (the
    margins and z ghost planes reach far outside the array): bit-identical to the oracle."""
    for (nx, ny, nz) in ((1, 1, 1), (2, 3, 1), (3, 1, 2), (5, 4, 3)):
        for prec, dt, dtype in (("dp", np.float64, gpu.GR_F64), ("sp", np.float32, gpu.GR_F32)):
            print("fused tiny", kernel, depth, (nx, ny, nz), prec, flush=True)
            u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 70 + nx + 10 * kernel)
            napply = 2 * depth + (0 if kernel == 0 else 1)
            with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel)) as st:
                st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
                st.set_config(zchunk=1 if nz > 1 else 0, t_block=depth)
                st.step(napply)
                assert st.stats()["last_launches"] < napply
                U1, U2 = st.download(_pad(u).shape)
            for a in range(napply):
                if a % 2 == 0:
                    orc.sweep(kernel, coef, u, v, roc2)
                else:
                    orc.sweep(kernel, coef, v, u, roc2)
            np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u, err_msg="%s %s" % ((nx, ny, nz), prec))
            np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v, err_msg="%s %s" % ((nx, ny, nz), prec))


@pytest.mark.parametrize("depth", [2, 4, 6])
def test_fused_bf16_coefficients_tiny_grids_bitexact(depth):
    """The bf16 coefficient hand-off (two values to a register) on grids smaller than one
    tile and than the pipeline depth: bit-identical to the FP32 oracle on the rounded
    coefficients."""
    kernel = 5
    for (nx, ny, nz) in ((1, 1, 1), (2, 3, 1), (3, 1, 2), (5, 4, 3)):
        print("fused bf16 tiny", depth, (nx, ny, nz), flush=True)
        u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, np.float32, 90 + nx + 3 * ny)
        coef = _bf16_round(coef)
        napply = 2 * depth + 1
        with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_COEF_BF16) as st:
            st.upload(_pad(u), _pad(v), None, _pad_coef(coef, kernel))
            st.set_config(zchunk=1 if nz > 1 else 0, t_block=depth)
            st.step(napply)
            assert st.stats()["last_launches"] < napply
            U1, U2 = st.download(_pad(u).shape)
        for a in range(napply):
            if a % 2 == 0:
                orc.sweep(kernel, coef, u, v, roc2)
            else:
                orc.sweep(kernel, coef, v, u, roc2)
        np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u, err_msg=str((nx, ny, nz)))
        np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v, err_msg=str((nx, ny, nz)))


def test_fused_fast_fp32_equals_single_sweep_fast():
    """GR_FAST fused passes use the same arithmetic as GR_FAST single sweeps."""
    nx, ny, nz = 80, 36, 24
    kernel = 0
    u, v, _, coef = _random_problem(kernel, nx, ny, nz, np.float32, 77)
    outs = []
    for depth in (1, 2):
        with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_ROC2_SCALAR, roc2=0.12,
                         coef=list(coef[:5])) as st:
            st.upload(_pad(u), _pad(v))
            st.set_config(t_block=depth)
            st.step(9)
            outs.append(st.download(_pad(u).shape))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_unequal_halos_fall_back_to_single_sweeps():
    """Fused passes assume U1 and U2 share girih's halo; otherwise the stepper must
    quietly use single sweeps and stay exact."""
    kernel, nx, ny, nz = 1, 40, 20, 12
    u, v, _, coef = _random_problem(kernel, nx, ny, nz, np.float64, 3)
    v[0, :, :] = 1.5  # a halo plane of U2 differs from U1's
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=_consts(coef, kernel)) as st:
        st.upload(_pad(u), _pad(v))
        st.set_config(t_block=4)
        st.step(10)
        assert st.stats()["last_launches"] == 10
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, None, nt=10)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


def test_girih_init_fused_matches_golden_semantics():
    """girih's closed-form state, diamond-stepper semantics (nt-1 applications),
    run with fused passes: U1/U2 equal the oracle's."""
    kernel, nx, ny, nz, nt, t_dim = 1, 16, 32, 24, 9, 3
    u, v, _ = orc.init_ref(kernel, nx, ny, nz)
    coef = orc.init_coef(kernel, u.shape)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=_consts(coef, kernel)) as st:
        st.init_fields(0)
        st.set_config(t_block=4)
        ntr = st.run_ts(2, nt, t_dim)
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, None, nt=ntr - 2)
    orc.sweep(kernel, coef, u, v)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


def test_autotune_keeps_fields_and_may_pick_fused():
    kernel, nx, ny, nz = 5, 96, 64, 40
    u, v, _, coef = _random_problem(kernel, nx, ny, nz, np.float32, 9)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32) as st:
        st.upload(_pad(u), _pad(v), None, _pad_coef(coef, kernel))
        best = st.autotune(4)
        assert best["t_block"] in (1, 2, 4, 6)
        st.step(7)
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, None, nt=6)
    orc.sweep(kernel, coef, u, v)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


def test_set_config_rejects_unavailable_depth():
    for kernel, depth in ((1, 7), (5, 3), (0, 4), (2, 2)):
        with gpu.Stepper(kernel, 16, 16, 16) as st:
            with pytest.raises(gpu.GirihError) as ei:
                st.set_config(t_block=depth)
            assert ei.value.code == gpu.GR_EUNSUPPORTED


def test_set_xcd_order_accepts_three_orders():
    with gpu.Stepper(0, 16, 16, 16) as st:
        for order in (0, 1, 2):
            st.set_xcd_order(order)
            assert st.stats()["xcd_order"] == order
        for bad in (-1, 3):
            with pytest.raises(gpu.GirihError) as ei:
                st.set_xcd_order(bad)
            assert ei.value.code == gpu.GR_EINVAL
        assert st.stats()["xcd_order"] == 2  # a rejected order leaves the setting unchanged


@pytest.mark.parametrize("kernel,dtype,flags,depth,tile", [
    (5, gpu.GR_F32, 0, 4, [64, 16]), (5, gpu.GR_F32, gpu.GR_COEF_BF16, 4, [64, 32]),
    (5, gpu.GR_F32, gpu.GR_COEF_BF16, 6, [64, 16]), (5, gpu.GR_F32, 0, 2, [64, 32]),
    (5, gpu.GR_F64, 0, 2, [64, 16]), (5, gpu.GR_F64, 0, 4, [64, 8]), (1, gpu.GR_F32, 0, 2, [128, 32]),
    (1, gpu.GR_F64, 0, 3, [64, 16]),
    (0, gpu.GR_F32, gpu.GR_FAST | gpu.GR_ROC2_SCALAR, 2, [64, 16])])
def test_stats_report_the_fused_pass_tile(kernel, dtype, flags, depth, tile):
    with gpu.Stepper(kernel, 40, 40, 24, dtype=dtype, flags=flags) as st:
        st.set_config(13, 0, 1)
        s = st.stats()
        assert [s["tile_x"], s["tile_y"]] == [128, 16] and s["t_block"] == 1
        st.set_config(t_block=depth)
        s = st.stats()
        assert [s["tile_x"], s["tile_y"]] == tile and s["t_block"] == depth
