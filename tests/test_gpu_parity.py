"""Parity of the HIP stepper (called through the C ABI) with girih's reference.

* golden: girih's own SB / diamond steppers (captured in tests/golden) must be
  reproduced bit for bit -- U1 and U2, DP and SP, kernels 0-5, including the
  diamond stepper's nt rounding and its nt-1 applications;
* random: seeded random fields/coefficients on ragged sizes (nx not a multiple
  of the 64-wide tile, ny not a multiple of the tile height) against the CPU
  oracle, bit-exact;
* FAST FP32 (kernel 0): error against the FP64 oracle within a stated bound.
"""
import hashlib

import numpy as np
import pytest

import oracle as orc
from girih_amd import gpu

pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _pad(a, alignment=8):
    return orc.to_padded(a, None, alignment)


def _pad_coef(coef, kernel):
    if orc.NUM_COEF_ARRAYS[kernel] == 0:
        return None
    return np.ascontiguousarray(np.stack([_pad(c) for c in coef]))


def _consts(coef, kernel):
    return list(coef[:orc.RADIUS[kernel] + 1]) if orc.NUM_COEF_ARRAYS[kernel] == 0 else None


@pytest.mark.parametrize("idx", range(42))
def test_golden_reference_cases(golden, idx):
    if idx >= len(golden):
        pytest.skip("no such case")
    c = golden[idx]
    print("case", c["name"], flush=True)
    k = c["kernel"]
    dt = np.float64 if c["precision"] == "dp" else np.float32
    u, v, roc2 = orc.init_ref(k, c["nx"], c["ny"], c["nz"], dtype=dt)
    coef = orc.init_coef(k, u.shape, dtype=dt)
    dtype = gpu.GR_F64 if c["precision"] == "dp" else gpu.GR_F32
    with gpu.Stepper(k, c["nx"], c["ny"], c["nz"], dtype=dtype, coef=_consts(coef, k)) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, k))
        t_dim = int(c["flags"][c["flags"].index("--t-dim") + 1]) if "--t-dim" in c["flags"] else 1
        nt_applied = st.run_ts(c["ts"], c["nt_flag"], t_dim)
        assert nt_applied == c["nt_applied"]
        U1, U2 = st.download(_pad(u).shape)
    assert list(U1.shape[::-1]) == c["ldomain_shape"]
    assert _sha(U1) == c["sha256_u1"], c["name"]
    assert _sha(U2) == c["sha256_u2"], c["name"]


def _random_problem(kernel, nx, ny, nz, dt, seed):
    rng = np.random.default_rng(seed)
    shp = orc.shape_of(kernel, nx, ny, nz)
    R = orc.RADIUS[kernel]
    u = np.zeros(shp, dtype=dt)
    u[R:-R, R:-R, R:-R] = rng.uniform(0, 1, (nz, ny, nx))
    v = np.zeros(shp, dtype=dt)
    v[R:-R, R:-R, R:-R] = rng.uniform(0, 1, (nz, ny, nx))
    roc2 = None
    if kernel == 0:
        roc2 = rng.uniform(0.05, 0.2, shp).astype(dt)
        coef = np.zeros(11, dtype=dt)
        w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
        coef[0] = 3 * w[0]
        coef[1:5] = w[1:5]
    elif orc.NUM_COEF_ARRAYS[kernel] == 0:
        coef = np.zeros(11, dtype=dt)
        coef[0], coef[1] = 0.4, 0.1
    else:
        na = orc.NUM_COEF_ARRAYS[kernel]
        coef = np.empty((na,) + shp, dtype=dt)
        coef[0] = rng.uniform(0.3, 0.5, shp)
        coef[1:] = rng.uniform(0, 0.1, (na - 1,) + shp)
    return u, v, roc2, coef


RAGGED = {
    4: [(67, 21, 17), (130, 9, 10), (8, 40, 12)],
    1: [(70, 19, 13), (129, 33, 7), (5, 3, 4)],
}


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("strm", [0, 1])
def test_random_ragged_bitexact(kernel, prec, strm):
    """strm 1 forces the non-temporal kernel instantiations that problems larger than
    the caches (the benchmark sizes) run, at sizes the oracle checks in seconds."""
    dt = np.float64 if prec == "dp" else np.float32
    dtype = gpu.GR_F64 if prec == "dp" else gpu.GR_F32
    for n, (nx, ny, nz) in enumerate(RAGGED[orc.RADIUS[kernel]]):
        print("kernel", kernel, prec, strm, (nx, ny, nz), flush=True)
        u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 100 * kernel + n)
        with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel)) as st:
            st.set_streaming(strm)
            st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
            st.step(5)  # odd count: ends with U1 written
            U1, U2 = st.download(_pad(u).shape)
        orc.run(kernel, coef, u, v, roc2, nt=4)
        orc.sweep(kernel, coef, u, v, roc2)
        nnx = u.shape[2]
        np.testing.assert_array_equal(U1[:, :, :nnx], u)
        np.testing.assert_array_equal(U2[:, :, :nnx], v)


def _num_variants(st):
    """Tile variants the stepper accepts (gr_set_config rejects the first one past the end)."""
    n = 0
    while n < 64:
        try:
            st.set_config(n, -1, -1)
        except gpu.GirihError:
            break
        n += 1
    st.set_config(0, -1, -1)
    return n


@pytest.mark.parametrize("kernel", [0, 1, 5])
@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("strm", [0, 1])
def test_every_tile_variant_bitexact(kernel, prec, strm):
    """Every launch configuration the autotuner may pick (tile variant, XCD order, z
    chunk, normal or non-temporal accesses) gives the oracle's result bit for bit on a
    ragged grid."""
    dt = np.float64 if prec == "dp" else np.float32
    dtype = gpu.GR_F64 if prec == "dp" else gpu.GR_F32
    nx, ny, nz = (290, 37, 13) if orc.RADIUS[kernel] == 4 else (300, 45, 11)
    u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, dt, 900 + kernel)
    uo, vo = u.copy(), v.copy()
    orc.run(kernel, coef, uo, vo, roc2, nt=2)
    orc.sweep(kernel, coef, uo, vo, roc2)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype, coef=_consts(coef, kernel)) as st:
        nvar = _num_variants(st)
        assert nvar == (17 if kernel == 0 else 16 if kernel in (1, 5) else 1)
        st.set_streaming(strm)
        for var in range(nvar):
            for order, zc in ((0, 0), (1, 5), (2, 3)):
                print("variant", kernel, prec, strm, var, order, zc, flush=True)
                st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
                st.set_config(var, zc, 1)
                st.set_xcd_order(order)
                st.step(3)
                U1, U2 = st.download(_pad(u).shape)
                nnx = u.shape[2]
                np.testing.assert_array_equal(U1[:, :, :nnx], uo, err_msg="variant %d" % var)
                np.testing.assert_array_equal(U2[:, :, :nnx], vo, err_msg="variant %d" % var)


@pytest.mark.parametrize("kernel", [0, 1, 5])
def test_tiny_grids_every_variant_bitexact(kernel):
    """Grids smaller than every tile, with single-cell dimensions and one-plane z chunks,
    in every tile variant and workgroup order: the clamped loads, the operand prefetch and
    the workgroup maps must still give the oracle's result bit for bit."""
    for (nx, ny, nz) in ((1, 1, 1), (2, 3, 1), (3, 1, 2), (1, 5, 3)):
        u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, np.float64, 300 + nx + 7 * ny + kernel)
        uo, vo = u.copy(), v.copy()
        orc.run(kernel, coef, uo, vo, roc2, nt=2)
        orc.sweep(kernel, coef, uo, vo, roc2)
        with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=_consts(coef, kernel)) as st:
            for var in range(_num_variants(st)):
                for order, zc in ((0, 0), (2, 1)):
                    print("tiny", kernel, (nx, ny, nz), var, order, zc, flush=True)
                    st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
                    st.set_config(var, zc, 1)
                    st.set_xcd_order(order)
                    st.step(3)
                    U1, U2 = st.download(_pad(u).shape)
                    nnx = u.shape[2]
                    np.testing.assert_array_equal(U1[:, :, :nnx], uo, err_msg="variant %d" % var)
                    np.testing.assert_array_equal(U2[:, :, :nnx], vo, err_msg="variant %d" % var)


WAVE_W = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
WAVE_COEF = [3 * WAVE_W[0]] + WAVE_W[1:5]


@pytest.mark.parametrize("roc2_kind", ["scalar", "field", "bf16"])
@pytest.mark.parametrize("strm", [0, 1])
def test_fast_every_variant_bitexact_vs_oracle(roc2_kind, strm):
    """The GR_FAST FP32 kernel-0 sweeps -- the benchmark's arithmetic -- for every tile
    variant, workgroup order, z chunk and cache policy, with a point source, equal the
    oracle's restatement of the GR_FAST arithmetic bit for bit.  strm 1 with the scalar
    roc2 is the instantiation the 1024^3 benchmark times."""
    nx, ny, nz = 290, 37, 13
    kernel = 0
    rng = np.random.default_rng(4242)
    shp = orc.shape_of(kernel, nx, ny, nz)
    u = np.zeros(shp, np.float32)
    v = np.zeros(shp, np.float32)
    u[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    v[4:-4, 4:-4, 4:-4] = rng.uniform(-1, 1, (nz, ny, nx))
    roc2 = None
    flags = gpu.GR_FAST
    if roc2_kind == "scalar":
        flags |= gpu.GR_ROC2_SCALAR
    else:
        roc2 = rng.uniform(0.05, 0.2, shp).astype(np.float32)
        if roc2_kind == "bf16":
            roc2 = _bf16_round(roc2)
            flags |= gpu.GR_COEF_BF16
    src = (np.cos(np.arange(4) * 1.3) * 0.5).astype(np.float64)
    pt = (4 + 131, 4 + 17, 4 + 6)
    uo, vo = u.copy(), v.copy()
    orc.run_fast(WAVE_COEF, uo, vo, roc2=roc2, roc2s=0.13, nt=3, src_idx=pt, src=src.astype(np.float32))
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=flags, roc2=0.13, coef=WAVE_COEF) as st:
        st.set_streaming(strm)
        for var in range(_num_variants(st)):
            for order, zc in ((0, 0), (1, 5), (2, 3)):
                print("fast variant", roc2_kind, strm, var, order, zc, flush=True)
                st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None)
                st.set_source(*pt, src)
                st.set_config(var, zc, 1)
                st.set_xcd_order(order)
                st.step(4)
                U1, U2 = st.download(_pad(u).shape)
                np.testing.assert_array_equal(U1[:, :, :shp[2]], uo, err_msg="variant %d" % var)
                np.testing.assert_array_equal(U2[:, :, :shp[2]], vo, err_msg="variant %d" % var)


def test_every_tile_variant_fast_identical():
    """GR_FAST arithmetic is per cell: every tile variant gives the same bits."""
    nx, ny, nz = 290, 37, 13
    u, v, _, coef = _random_problem(0, nx, ny, nz, np.float32, 77)
    ref = None
    with gpu.Stepper(0, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_ROC2_SCALAR, roc2=0.13,
                     coef=list(coef[:5])) as st:
        for var in range(_num_variants(st)):
            st.upload(_pad(u), _pad(v))
            st.set_config(var, 7 if var % 2 else 0, 1)
            st.step(4)
            out = st.download(_pad(u).shape)
            if ref is None:
                ref = out
            np.testing.assert_array_equal(out[0], ref[0], err_msg="variant %d" % var)
            np.testing.assert_array_equal(out[1], ref[1], err_msg="variant %d" % var)


def _bf16_round(a):
    """float32 array rounded to bfloat16 values (nearest even), kept as float32."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("kernel", [0, 2, 3, 4, 5])
def test_bf16_coefficient_fields_bitexact(kernel):
    """GR_COEF_BF16 stores roc2 / coefficient arrays as bfloat16 and computes in
    FP32 in girih's operand order: bit-identical to the FP32 oracle run on the same
    bf16-rounded coefficients."""
    nx, ny, nz = (70, 21, 13) if orc.RADIUS[kernel] == 4 else (77, 35, 12)
    u, v, roc2, coef = _random_problem(kernel, nx, ny, nz, np.float32, 300 + kernel)
    if roc2 is not None:
        roc2 = _bf16_round(roc2)
    if orc.NUM_COEF_ARRAYS[kernel]:
        coef = _bf16_round(coef)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_COEF_BF16, coef=_consts(coef, kernel)) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2) if roc2 is not None else None, _pad_coef(coef, kernel))
        st.step(5)
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, roc2, nt=4)
    orc.sweep(kernel, coef, u, v, roc2)
    nnx = u.shape[2]
    np.testing.assert_array_equal(U1[:, :, :nnx], u)
    np.testing.assert_array_equal(U2[:, :, :nnx], v)


def test_device_random_coefficients_do_not_depend_on_precision():
    """gr_init_coef_uniform and gr_init_fields(mode 1) draw from logical cell indices:
    an FP32 and an FP64 problem with the same seeds hold the same values up to
    rounding, so their runs differ only by FP32 arithmetic (and the fields a new
    problem is created with are zero before its first kernel runs)."""
    # 256^3: large enough that zero-filling a new problem takes longer than its first
    # kernels (the fills must be ordered with the handle's stream, not the null stream)
    for kernel, n in ((5, (70, 33, 21)), (3, (40, 20, 18)), (4, (36, 20, 18)), (5, (256, 256, 256))):
        hs = {}
        for dt in (gpu.GR_F64, gpu.GR_F32):
            st = gpu.Stepper(kernel, *n, dtype=dt)
            st.init_coef_uniform(5, 0.3, 0.5, 0.0, 0.1)
            st.init_fields(1, 12345)
            st.step(12)
            hs[dt] = st
        c = hs[gpu.GR_F32].compare(1, hs[gpu.GR_F64], 1)
        print("kernel", kernel, "FP32 vs FP64 rel L2", c["rel_l2"], flush=True)
        assert c["sq_ref"] > 0 and c["rel_l2"] < 1e-5
        for st in hs.values():
            st.close()


@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_device_closed_form_init_matches_girih(prec):
    """gr_init_fields(mode 0) reproduces girih's closed-form state bit for bit."""
    dt = np.float64 if prec == "dp" else np.float32
    dtype = gpu.GR_F64 if prec == "dp" else gpu.GR_F32
    for kernel, (nx, ny, nz) in [(0, (37, 20, 11)), (1, (64, 9, 17)), (5, (13, 9, 11))]:
        u, v, roc2 = orc.init_ref(kernel, nx, ny, nz, dtype=dt)
        with gpu.Stepper(kernel, nx, ny, nz, dtype=dtype) as st:
            st.init_fields(0)
            U1, U2 = st.download(_pad(u).shape)
        np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
        np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


def test_source_injection_matches_oracle():
    """nb_naive_ts.c:190,196 point source, k0 FP64 from a zero field."""
    nx, ny, nz = 40, 24, 20
    kernel = 0
    shp = orc.shape_of(kernel, nx, ny, nz)
    u = np.zeros(shp)
    v = np.zeros(shp)
    roc2 = np.full(shp, 0.16)
    coef = np.zeros(11)
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef[0] = 3 * w[0]
    coef[1:5] = w[1:5]
    src = np.sin(np.arange(10) * 0.7) + 0.1
    pt = (4 + 17, 4 + 11, 4 + 9)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F64, coef=list(coef[:5])) as st:
        st.upload(_pad(u), _pad(v), _pad(roc2))
        st.set_source(*pt, src)
        st.step(10)
        U1, U2 = st.download(_pad(u).shape)
    orc.run(kernel, coef, u, v, roc2, nt=10, src_idx=pt, src=src)
    np.testing.assert_array_equal(U1[:, :, :u.shape[2]], u)
    np.testing.assert_array_equal(U2[:, :, :u.shape[2]], v)


def test_fast_fp32_error_vs_fp64_oracle():
    """GR_FAST (neighbour-difference pair sums, FMA, compensated update) stays within
    1e-6 relative L2 of the FP64 oracle after 40 steps of a smooth wave problem (the
    plain reference-order FP32 form gives ~9e-6 here, dominated by cancellation)."""
    nx, ny, nz = 72, 40, 36
    kernel = 0
    shp = orc.shape_of(kernel, nx, ny, nz)
    kk, jj, ii = np.meshgrid(*[np.arange(n) for n in shp], indexing="ij")
    g = np.exp(-((ii - shp[2] / 2) ** 2 + (jj - shp[1] / 2) ** 2 + (kk - shp[0] / 2) ** 2) / 30.0)
    g[:4] = 0
    g[-4:] = 0
    g[:, :4] = 0
    g[:, -4:] = 0
    g[:, :, :4] = 0
    g[:, :, -4:] = 0
    g = g.astype(np.float32).astype(np.float64)  # same (FP32-representable) start for both
    coef = np.zeros(11)
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef[0] = 3 * w[0]
    coef[1:5] = w[1:5]
    u64, v64 = g.copy(), g.copy()
    roc2 = np.full(shp, 0.16)
    orc.run(kernel, coef, u64, v64, roc2, nt=40)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_FAST | gpu.GR_ROC2_SCALAR, roc2=0.16,
                     coef=list(coef[:5])) as st:
        g32 = _pad(g.astype(np.float32))
        st.upload(g32, g32.copy())
        st.step(40)
        U1, U2 = st.download(g32.shape)
    with gpu.Stepper(kernel, nx, ny, nz, dtype=gpu.GR_F32, flags=gpu.GR_ROC2_SCALAR, roc2=0.16,
                     coef=list(coef[:5])) as st:
        st.upload(g32, g32.copy())
        st.step(40)
        _, P2 = st.download(g32.shape)
    ref = v64[4:-4, 4:-4, 4:-4]
    got = U2[4:-4, 4:-4, 4:4 + nx].astype(np.float64)
    plain = P2[4:-4, 4:-4, 4:4 + nx].astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    rel_plain = np.linalg.norm(plain - ref) / np.linalg.norm(ref)
    print("FP32 rel L2 vs FP64 after 40 steps: FAST %.3e, reference order %.3e" % (rel, rel_plain))
    assert rel < 1e-6
    assert rel < rel_plain


def test_level2_sweep_dev_on_torch_memory():
    """gr_sweep_dev (KERNEL_SIG with device pointers) on girih-layout device arrays,
    for a sub-box, against the oracle applied to the same box."""
    torch = pytest.importorskip("torch")
    kernel = 5
    nx, ny, nz = 30, 14, 9
    u, v, _, coef = _random_problem(kernel, nx, ny, nz, np.float64, 5)
    pu, pv, pc = _pad(u), _pad(v), _pad_coef(coef, kernel)
    du = torch.from_numpy(pu).cuda()
    dv = torch.from_numpy(pv).cuda()
    dc = torch.from_numpy(pc).cuda()
    shape = (pu.shape[2], pu.shape[1], pu.shape[0])
    box = (1, 1, 1, nx + 1, ny + 1, nz + 1)
    gpu.sweep_dev(kernel, gpu.GR_F64, shape, box, dc.data_ptr(), du.data_ptr(), dv.data_ptr())
    out = du.cpu().numpy()
    orc.sweep(kernel, coef, u, v)
    np.testing.assert_array_equal(out[:, :, :u.shape[2]], u)
    # a sub-box leaves the rest untouched
    du2 = torch.from_numpy(_pad(np.zeros_like(u))).cuda()
    gpu.sweep_dev(kernel, gpu.GR_F64, shape, (3, 2, 4, 20, 9, 7), dc.data_ptr(), du2.data_ptr(), dv.data_ptr())
    out2 = du2.cpu().numpy()[:, :, :u.shape[2]]
    np.testing.assert_array_equal(out2[4:7, 2:9, 3:20], u[4:7, 2:9, 3:20])
    out2[4:7, 2:9, 3:20] = 0
    assert not np.any(out2)
