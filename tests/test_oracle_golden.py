"""Pin the CPU oracle against girih's own steppers (golden vectors) and against the
mathematics where the reference leaves something open.

The golden vectors in tests/golden/ were captured from girih's C sources compiled
where they lie (oracle/Makefile `ref`, oracle/ref_dump.c); every captured run also
passed girih's own exact --verify comparison.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _oracle_case(c):
    """Oracle state equivalent to the reference stepper's final (U1, U2).

    Spatial blocking (ts 0) applies 2*ceil(nt/2) steps: U1 = step N-1, U2 = step N.
    The diamond stepper (ts 2) applies one step less than its rounded nt:
    U1 = step nt-1 (the buffer girih verifies), U2 = step nt-2 -- i.e. the oracle
    pair after nt-2 steps followed by one half step U1 <- f(U2)."""
    dt = np.float64 if c["precision"] == "dp" else np.float32
    k = c["kernel"]
    u, v, roc2 = orc.init_ref(k, c["nx"], c["ny"], c["nz"], dtype=dt)
    coef = orc.init_coef(k, u.shape, dtype=dt)
    if c["ts"] == 2:
        n = orc.run(k, coef, u, v, roc2, nt=c["nt_applied"] - 2)
        orc.sweep(k, coef, u, v, roc2)
    else:
        n = orc.run(k, coef, u, v, roc2, nt=c["nt_applied"])
        assert n == 2 * ((c["nt_applied"] + 1) // 2)
    return u, v


@pytest.mark.parametrize("idx", range(42))
def test_oracle_matches_reference_golden(golden, idx):
    """Bit-exact: oracle U1/U2 (placed in girih's padded layout) hash-equal to the
    reference stepper's U1/U2 (SB and diamond steppers, DP and SP, kernels 0-5)."""
    if idx >= len(golden):
        pytest.skip("no such case")
    c = golden[idx]
    u, v = _oracle_case(c)
    pu, pv = orc.to_padded(u, c["kernel"]), orc.to_padded(v, c["kernel"])
    assert list(pu.shape[::-1]) == c["ldomain_shape"]
    assert _sha(pu) == c["sha256_u1"], c["name"]
    assert _sha(pv) == c["sha256_u2"], c["name"]


def test_golden_full_arrays():
    """Full U1/U2 arrays for the smallest cases compare element-exactly."""
    import json
    arrays = np.load(os.path.join(REPO, "tests", "golden", "ref_arrays.npz"))
    with open(os.path.join(REPO, "tests", "golden", "ref_cases.json")) as f:
        cases = {c["name"]: c for c in json.load(f)["cases"]}
    names = sorted({k.split("__")[0] for k in arrays.files})
    assert len(names) >= 6
    for name in names:
        c = cases[name]
        u, v = _oracle_case(c)
        ref_u1 = arrays[name + "__u1"]
        ref_u2 = arrays[name + "__u2"]
        nnz, nny, nnx = u.shape
        np.testing.assert_array_equal(ref_u1[:, :, :nnx], u)
        np.testing.assert_array_equal(ref_u2[:, :, :nnx], v)
        assert not np.any(ref_u1[:, :, nnx:]), "reference padding must stay zero"


def test_mwd_nt_rounding_matches_reference(golden):
    for c in golden:
        if c["ts"] == 2:
            td = int(c["flags"][c["flags"].index("--t-dim") + 1])
            assert orc.mwd_round_nt(c["nt_flag"], td) == c["nt_applied"], c["name"]
    # diamond_utils.c:1042-1056 worked examples (100 -> 106 with t_dim 3 was observed)
    assert orc.mwd_round_nt(100, 3) == 106
    assert orc.mwd_round_nt(10, 3) == 10
    assert orc.mwd_round_nt(1000, 1) == 1002


def test_reference_live_if_built(tmp_path):
    """When the reference build exists (this container), compare a fresh run with
    different sizes than the fixtures (k0 SB, k5 1WD)."""
    import subprocess
    exe = os.path.join(REPO, "oracle", "_ref", "ref_dump_dp")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built")
    for (k, ts, extra, nx, ny, nz, nt) in [
        (0, 0, ["--thread-group-size", "1"], 12, 20, 18, 6),
        (5, 2, ["--thread-group-size", "1", "--thx", "1", "--thy", "1", "--thz", "1", "--t-dim", "1",
                "--num-wavefronts", "1"], 10, 16, 12, 8),
    ]:
        out = str(tmp_path / "d.bin")
        flags = ["--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--nt", str(nt), "--target-ts", str(ts),
                 "--target-kernel", str(k), "--verbose", "0"] + extra
        r = subprocess.run([exe, out] + flags, env=dict(os.environ, OMP_NUM_THREADS="1"),
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout + r.stderr
        raw = open(out, "rb").read()
        hdr = np.frombuffer(raw[:48], dtype=np.int32)
        nnx, nny, nnz = hdr[6], hdr[7], hdr[8]
        n = int(nnx) * int(nny) * int(nnz)
        u1 = np.frombuffer(raw[48:48 + 8 * n], dtype=np.float64).reshape(nnz, nny, nnx)
        u, v, roc2 = orc.init_ref(k, nx, ny, nz)
        coef = orc.init_coef(k, u.shape)
        orc.run(k, coef, u, v, roc2, nt=int(hdr[9]))
        # U1 is the buffer girih verifies: step N-1 for both steppers
        np.testing.assert_array_equal(u1[:, :, :u.shape[2]], u)


# ----------------------------------------------------------------------------
# Mathematical pins (independent of the reference binary)
# ----------------------------------------------------------------------------

def test_k1_dirichlet_sine_eigenmode():
    """k1 with zero halo: sin(pi a i/(n+1)) sin(pi b j/(m+1)) sin(pi c k/(l+1)) is an
    exact eigenvector of the discrete operator, with eigenvalue
    c0 + 2 c1 (cos(pi a/(n+1)) + cos(pi b/(m+1)) + cos(pi c/(l+1)))."""
    nx, ny, nz = 14, 10, 12
    a, b, c = 2, 3, 1
    i = np.arange(nx + 2)
    j = np.arange(ny + 2)
    k = np.arange(nz + 2)
    f = (np.sin(np.pi * a * k[:, None, None] * 0 + np.pi * c * k[:, None, None] / (nz + 1))
         * np.sin(np.pi * b * j[None, :, None] / (ny + 1)) * np.sin(np.pi * a * i[None, None, :] / (nx + 1)))
    f = np.ascontiguousarray(f)
    coef = np.zeros(11)
    coef[0], coef[1] = 0.4, 0.1
    u = np.zeros_like(f)
    orc.sweep(1, coef, u, f)
    lam = 0.4 + 0.2 * (np.cos(np.pi * a / (nx + 1)) + np.cos(np.pi * b / (ny + 1)) + np.cos(np.pi * c / (nz + 1)))
    np.testing.assert_allclose(u[1:-1, 1:-1, 1:-1], lam * f[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-14)


def test_k0_quadratic_laplacian_is_exact():
    """With the 8th-order FD weights (3-D centre 3*w0), Lap8 of a quadratic field is exact:
    x^2+2y^2+3z^2 -> 2+4+6 = 12.  One k0 step with u^{n-1} = u^n gives
    u^{n+1} = u^n + roc2 * 12 (checks every tap, sign and axis of the 25-point star)."""
    w = [-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560]
    coef = np.zeros(11)
    coef[0] = 3 * w[0]
    coef[1:5] = w[1:5]
    nx, ny, nz = 9, 11, 10
    kk, jj, ii = np.meshgrid(np.arange(nz + 8), np.arange(ny + 8), np.arange(nx + 8), indexing="ij")
    q = (ii - 7.0) ** 2 + 2 * (jj - 5.0) ** 2 + 3 * (kk - 6.0) ** 2
    v = np.ascontiguousarray(q)
    u = v.copy()
    roc2 = np.full_like(v, 0.16)
    orc.sweep(0, coef, u, v, roc2)
    np.testing.assert_allclose(u[4:-4, 4:-4, 4:-4], q[4:-4, 4:-4, 4:-4] + 0.16 * 12, rtol=1e-12, atol=1e-9)


def _brute_k0(coef, u, v, roc2):
    out = u.copy()
    nnz, nny, nnx = v.shape
    for k in range(4, nnz - 4):
        for j in range(4, nny - 4):
            for i in range(4, nnx - 4):
                lap = coef[0] * v[k, j, i]
                for r in range(1, 5):
                    lap += coef[r] * (v[k, j, i + r] + v[k, j, i - r] + v[k, j + r, i] + v[k, j - r, i]
                                      + v[k + r, j, i] + v[k - r, j, i])
                out[k, j, i] = 2 * v[k, j, i] - u[k, j, i] + roc2[k, j, i] * lap
    return out


def test_k0_brute_force_long_double_tiny():
    """Tiny random grid: oracle k0 equals an independent brute-force evaluation in
    long double to within FP64 rounding of the 25 terms."""
    rng = np.random.default_rng(7)
    shp = (10, 11, 12)
    v = rng.standard_normal(shp)
    u = rng.standard_normal(shp)
    roc2 = rng.uniform(0.1, 0.2, shp)
    coef = orc.default_coef()
    ref = _brute_k0(coef.astype(np.longdouble), u.astype(np.longdouble), v.astype(np.longdouble),
                    roc2.astype(np.longdouble))
    out = u.copy()
    orc.sweep(0, coef, out, v, roc2)
    np.testing.assert_allclose(out, ref.astype(np.float64), rtol=0, atol=1e-13)


@pytest.mark.parametrize("kernel", [1, 2, 3, 4, 5])
def test_first_order_kernels_linear(kernel):
    """Linearity f(a x + b y) = a f(x) + b f(y) of the first-order kernels, and the
    variable-coefficient kernels agree with explicit neighbour sums on random coefs."""
    rng = np.random.default_rng(kernel)
    R = orc.RADIUS[kernel]
    shp = (7 + 2 * R, 6 + 2 * R, 5 + 2 * R)
    x = rng.standard_normal(shp)
    y = rng.standard_normal(shp)
    na = orc.NUM_COEF_ARRAYS[kernel]
    coef = rng.uniform(-1, 1, 11) if na == 0 else rng.uniform(-1, 1, (na,) + shp)
    fx, fy, fz = np.zeros(shp), np.zeros(shp), np.zeros(shp)
    orc.sweep(kernel, coef, fx, x)
    orc.sweep(kernel, coef, fy, y)
    orc.sweep(kernel, coef, fz, np.ascontiguousarray(2.5 * x - 0.5 * y))
    np.testing.assert_allclose(fz, 2.5 * fx - 0.5 * fy, rtol=1e-12, atol=1e-12)
    # explicit neighbour sums per kernel definition (stencils.c:73-201)
    s = (slice(R, -R),) * 3

    def sh(a, dk, dj, di):
        nz, ny, nx = a.shape
        return a[R + dk:nz - R + dk, R + dj:ny - R + dj, R + di:nx - R + di]
    C = (lambda m: coef[m][s]) if na else None
    if kernel == 5:
        exp = (C(0) * sh(x, 0, 0, 0) + C(1) * sh(x, 0, 0, -1) + C(2) * sh(x, 0, 0, 1) + C(3) * sh(x, 0, -1, 0)
               + C(4) * sh(x, 0, 1, 0) + C(5) * sh(x, -1, 0, 0) + C(6) * sh(x, 1, 0, 0))
    elif kernel == 1:
        exp = coef[0] * sh(x, 0, 0, 0) + coef[1] * sum(sh(x, *d) for d in
                                                      [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)])
    elif kernel == 2:
        exp = C(0) * sh(x, 0, 0, 0) + C(1) * sum(sh(x, *d) for d in
                                               [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)])
    elif kernel == 3:
        exp = (C(0) * sh(x, 0, 0, 0) + C(1) * (sh(x, 0, 0, 1) + sh(x, 0, 0, -1))
               + C(2) * (sh(x, 0, 1, 0) + sh(x, 0, -1, 0)) + C(3) * (sh(x, 1, 0, 0) + sh(x, -1, 0, 0)))
    else:  # kernel 4: COEF(1+3(r-1)+axis) for axis x, y, z
        exp = C(0) * sh(x, 0, 0, 0)
        for r in range(1, 5):
            exp = exp + C(1 + 3 * (r - 1)) * (sh(x, 0, 0, r) + sh(x, 0, 0, -r))
            exp = exp + C(2 + 3 * (r - 1)) * (sh(x, 0, r, 0) + sh(x, 0, -r, 0))
            exp = exp + C(3 + 3 * (r - 1)) * (sh(x, r, 0, 0) + sh(x, -r, 0, 0))
    np.testing.assert_allclose(fx[s], exp, rtol=1e-12, atol=1e-12)


def test_source_injection_semantics():
    """nb_naive_ts.c:190,196: u gets src[it] after the first half step, v gets src[it+1]
    after the second.  With a zero field and k1 the response is linear in the source,
    and a delta at step 0 reaches its neighbours with weight c1 one step later."""
    nx = ny = nz = 6
    u = np.zeros((8, 8, 8))
    v = np.zeros_like(u)
    coef = np.zeros(11)
    coef[0], coef[1] = 0.4, 0.1
    src = np.array([1.0, 0.0, 0.0, 0.0])
    orc.run(1, coef, u, v, None, nt=2, src_idx=(3, 3, 3), src=src)
    # step 1: u = f(v=0) = 0, then u[3,3,3] += 1; step 2: v = f(u) -> v[3,3,3]=0.4, neighbours 0.1
    assert u[3, 3, 3] == 1.0 and np.count_nonzero(u) == 1
    assert v[3, 3, 3] == 0.4 and v[3, 3, 4] == 0.1 and v[4, 3, 3] == 0.1 and np.count_nonzero(v) == 7
    # linearity in the source
    u2 = np.zeros_like(u)
    v2 = np.zeros_like(u)
    orc.run(1, coef, u2, v2, None, nt=4, src_idx=(2, 3, 4), src=np.array([0.5, -1.0, 2.0, 0.25]))
    u3 = np.zeros_like(u)
    v3 = np.zeros_like(u)
    orc.run(1, coef, u3, v3, None, nt=4, src_idx=(2, 3, 4), src=np.array([1.0, -2.0, 4.0, 0.5]))
    np.testing.assert_allclose(2 * v2, v3, rtol=1e-15)


def test_compare_detects_single_ulp():
    """compare_results_std semantics: exact; one flipped ulp is a failure."""
    u, v, _ = orc.init_ref(1, 6, 5, 4)
    R = 1
    target = np.ascontiguousarray(u[R:-R, R:-R, R:-R])
    ok, emax, el1 = orc.compare(u, target, R)
    assert ok and emax == 0 and el1 == 0
    t2 = target.copy()
    t2[1, 2, 3] = np.nextafter(t2[1, 2, 3], 10)
    ok, emax, el1 = orc.compare(u, t2, R)
    assert not ok and emax > 0
