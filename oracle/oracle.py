"""ctypes binding of the CPU oracle (lib/libgirih_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C restatement of girih's serial
verification reference (src/verification.c:52-479, 823-860).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module;
the HIP product path never calls it.

Arrays are numpy arrays in the unpadded halo-inclusive layout of
verification.c:61-64, indexed [k, j, i] with shape (nz+2R, ny+2R, nx+2R).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgirih_oracle.so")

RADIUS = {0: 4, 1: 1, 2: 1, 3: 1, 4: 4, 5: 1}
TIME_ORDER = {0: 2, 1: 1, 2: 1, 3: 1, 4: 1, 5: 1}
NUM_COEF_ARRAYS = {0: 0, 1: 0, 2: 2, 3: 4, 4: 13, 5: 7}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle oracle` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        for sfx in ("f64", "f32"):
            f = getattr(L, "gro_sweep_" + sfx)
            f.argtypes = [I, P, P, P, P, P]
            f.restype = I
            f = getattr(L, "gro_run_" + sfx)
            f.argtypes = [I, P, P, P, P, P, I, ctypes.c_int64, P]
            f.restype = I
            f = getattr(L, "gro_init_ref_" + sfx)
            f.argtypes = [I, I, I, I, P, P, P]
            f.restype = None
            f = getattr(L, "gro_init_coef_" + sfx)
            f.argtypes = [I, ctypes.c_uint64, P, P]
            f.restype = None
            f = getattr(L, "gro_compare_" + sfx)
            f.argtypes = [P, P, I, I, I, I, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)]
            f.restype = I
            f = getattr(L, "gro_default_coef_" + sfx)
            f.argtypes = [P]
            f.restype = None
        L.gro_sweep_fast_f32.argtypes = [P, P, P, P, P, ctypes.c_double]
        L.gro_sweep_fast_f32.restype = None
        L.gro_run_fast_f32.argtypes = [P, P, P, P, P, ctypes.c_double, I, ctypes.c_int64, P]
        L.gro_run_fast_f32.restype = I
        L.gro_mwd_round_nt.argtypes = [I, I]
        L.gro_mwd_round_nt.restype = I
        L.gro_shot_f64.argtypes = [P, P, P, I, ctypes.c_int64, P, I, P, P, P, ctypes.c_double]
        L.gro_shot_f64.restype = I
        _lib = L
    return _lib


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError("oracle supports float64 and float32, got %s" % dtype)


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def default_coef(dtype=np.float64):
    out = np.zeros(11, dtype=dtype)
    getattr(lib(), "gro_default_coef_" + _sfx(dtype))(_ptr(out))
    return out


def mwd_round_nt(nt, t_dim):
    return lib().gro_mwd_round_nt(int(nt), int(t_dim))


def shape_of(kernel, nx, ny, nz):
    R = RADIUS[kernel]
    return (nz + 2 * R, ny + 2 * R, nx + 2 * R)


def init_ref(kernel, nx, ny, nz, dtype=np.float64):
    """Closed-form initial u, v, roc2 of verification.c:219-251."""
    shp = shape_of(kernel, nx, ny, nz)
    u = np.empty(shp, dtype=dtype)
    v = np.empty(shp, dtype=dtype)
    roc2 = np.empty(shp, dtype=dtype) if TIME_ORDER[kernel] == 2 else None
    getattr(lib(), "gro_init_ref_" + _sfx(dtype))(kernel, nx, ny, nz, _ptr(u), _ptr(v), _ptr(roc2))
    return u, v, roc2


def init_coef(kernel, shape, g_coef=None, dtype=np.float64):
    """Coefficients of verification.c:92-216 (constant list or SoA arrays)."""
    if g_coef is None:
        g_coef = default_coef(dtype)
    g_coef = np.ascontiguousarray(g_coef, dtype=dtype)
    n_domain = int(np.prod(shape))
    na = NUM_COEF_ARRAYS[kernel]
    coef = np.zeros(11, dtype=dtype) if na == 0 else np.empty((na,) + tuple(shape), dtype=dtype)
    getattr(lib(), "gro_init_coef_" + _sfx(dtype))(kernel, n_domain, _ptr(g_coef), _ptr(coef))
    return coef


def sweep(kernel, coef, u, v, roc2=None):
    """u <- f(v) once over the interior (in place on u)."""
    shape = (ctypes.c_int * 3)(u.shape[2], u.shape[1], u.shape[0])
    r = getattr(lib(), "gro_sweep_" + _sfx(u.dtype))(kernel, shape, _ptr(coef), _ptr(u), _ptr(v), _ptr(roc2))
    if r != 0:
        raise ValueError("bad kernel id %d" % kernel)


def run(kernel, coef, u, v, roc2=None, nt=2, src_idx=None, src=None):
    """Time loop of verification.c:281-284 (in place on u and v); returns applied steps.

    src_idx: (i, j, k) halo-inclusive point for the nb_naive_ts.c:190,196 source."""
    shape = (ctypes.c_int * 3)(u.shape[2], u.shape[1], u.shape[0])
    flat = -1
    if src_idx is not None:
        i, j, k = src_idx
        flat = (k * u.shape[1] + j) * u.shape[2] + i
        src = np.ascontiguousarray(src, dtype=u.dtype)
        assert src.size >= 2 * ((nt + 1) // 2)
    n = getattr(lib(), "gro_run_" + _sfx(u.dtype))(kernel, shape, _ptr(coef), _ptr(u), _ptr(v), _ptr(roc2),
                                                   int(nt), flat, _ptr(src))
    if n < 0:
        raise ValueError("bad kernel id %d" % kernel)
    return n


def sweep_fast(coef, u, v, roc2=None, roc2s=0.0):
    """FP32 kernel 0 in the library's GR_FAST arithmetic (u <- f(v) in place on u);
    coef holds c0..c4 (taken as float64 and rounded like the library does); roc2 is a
    float32 array or None for the scalar roc2s."""
    assert u.dtype == np.float32 and v.dtype == np.float32
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64)[:5])
    shape = (ctypes.c_int * 3)(u.shape[2], u.shape[1], u.shape[0])
    r = None if roc2 is None else np.ascontiguousarray(roc2, dtype=np.float32)
    lib().gro_sweep_fast_f32(shape, _ptr(c), _ptr(u), _ptr(v), _ptr(r), float(roc2s))


def run_fast(coef, u, v, roc2=None, roc2s=0.0, nt=2, src_idx=None, src=None):
    """run() for the GR_FAST sweep (FP32 kernel 0); returns applied steps."""
    assert u.dtype == np.float32 and v.dtype == np.float32
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64)[:5])
    shape = (ctypes.c_int * 3)(u.shape[2], u.shape[1], u.shape[0])
    r = None if roc2 is None else np.ascontiguousarray(roc2, dtype=np.float32)
    flat = -1
    if src_idx is not None:
        i, j, k = src_idx
        flat = (k * u.shape[1] + j) * u.shape[2] + i
        src = np.ascontiguousarray(src, dtype=np.float32)
        assert src.size >= 2 * ((nt + 1) // 2)
    return lib().gro_run_fast_f32(shape, _ptr(c), _ptr(u), _ptr(v), _ptr(r), float(roc2s), int(nt), flat,
                                  _ptr(src) if flat >= 0 else None)


def compare(u, target, R):
    """compare_results_std (verification.c:823-860); target is the nx*ny*nz interior."""
    nz, ny, nx = target.shape
    emax = ctypes.c_double()
    el1 = ctypes.c_double()
    ok = getattr(lib(), "gro_compare_" + _sfx(u.dtype))(_ptr(u), _ptr(np.ascontiguousarray(target)), nx, ny, nz, R,
                                                        ctypes.byref(emax), ctypes.byref(el1))
    return bool(ok), emax.value, el1.value


def to_padded(a, kernel, alignment=8):
    """Place an unpadded oracle array into girih's padded ldomain layout (utils.c:367-374):
    nnx rounded up to a multiple of `alignment` elements, pad cells zero."""
    nnz, nny, nnx = a.shape
    pad = (-nnx) % alignment
    out = np.zeros((nnz, nny, nnx + pad), dtype=a.dtype)
    out[:, :, :nnx] = a
    return out


def shot(coef, roc2, nt, src_idx, src, rec_k, res=None, dt_over_h=1.0):
    """Seismic shot on kernel 0 in FP64 (adjoint.c): zero initial fields, point
    source src[a] at halo-inclusive (i, j, k) = src_idx (or None), traces of plane
    rec_k.  Returns (trace [nt, ny, nx], grad [nz, ny, nx] or None); grad is the
    gradient of sum_a <res[a], trace[a]> with respect to m = dt_over_h**2 / roc2."""
    roc2 = np.ascontiguousarray(roc2, dtype=np.float64)
    nnz, nny, nnx = roc2.shape
    nx, ny, nz = nnx - 8, nny - 8, nnz - 8
    shape = (ctypes.c_int * 3)(nnx, nny, nnz)
    c = np.zeros(5)
    c[:] = np.asarray(coef, dtype=np.float64)[:5]
    flat = -1
    if src_idx is not None:
        i, j, k = src_idx
        flat = (k * nny + j) * nnx + i
        src = np.ascontiguousarray(src, dtype=np.float64)
        assert src.size >= nt
    trace = np.zeros((nt, ny, nx))
    grad = None
    if res is not None:
        res = np.ascontiguousarray(res, dtype=np.float64)
        assert res.shape == (nt, ny, nx)
        grad = np.zeros((nz, ny, nx))
    rc = lib().gro_shot_f64(shape, _ptr(c), _ptr(roc2), int(nt), flat, _ptr(src) if flat >= 0 else None, int(rec_k),
                            _ptr(trace), _ptr(res), _ptr(grad), float(dt_over_h))
    if rc != 0:
        raise ValueError("gro_shot_f64 failed (%d)" % rc)
    return trace, grad
