"""ctypes host binding of libgirih_gpu.so (the C ABI in include/girih_gpu.h).

The binding mirrors girih's stepper interface (src/wrappers.h:28-36,
src/data_structures.h:236-297): a problem is (kernel, precision, nx, ny, nz),
host fields use girih's padded layout (utils.c:367-374, x fastest, numpy shape
(nnz, nny, nnx_padded)), and a run follows girih's step pairing
(nb_naive_ts.c:187-203).  The library is loaded from girih_amd/lib; there is no
fallback: if it is missing, every call raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libgirih_gpu.so")

GR_OK, GR_EINVAL, GR_EHIP, GR_ENOMEM, GR_EUNSUPPORTED = 0, -1, -2, -3, -4
GR_F64, GR_F32 = 0, 1
GR_ROC2_SCALAR, GR_FAST, GR_COEF_BF16 = 0x1, 0x2, 0x4

RADIUS = {0: 4, 1: 1, 2: 1, 3: 1, 4: 4, 5: 1}
TIME_ORDER = {0: 2, 1: 1, 2: 1, 3: 1, 4: 1, 5: 1}
NUM_COEF_ARRAYS = {0: 0, 1: 0, 2: 2, 3: 4, 4: 13, 5: 7}
KERNEL_NAMES = {0: "iso_ref", 1: "iso_ref_2space_1time", 2: "iso_ref_2space_1time_var",
                3: "iso_ref_2space_1time_var_axsym", 4: "iso_ref_8space_1time_var_axsym",
                5: "iso_ref_2space_1time_var_nosym"}

# girih's default coefficient list (utils.c:116-119)
G_COEF = (-0.28472, 0.16000, -0.02000, 0.00254, -0.00018, -0.18472, 0.19, -0.0500, 0.00554, -0.0009, 0.00354)


class GirihError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("girih_gpu error %d: %s" % (code, msg))
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int), ("dtype", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("flags", ctypes.c_int), ("t_block", ctypes.c_int),
                ("reserved", ctypes.c_int), ("roc2", ctypes.c_double), ("coef", ctypes.c_double * 11)]


class Stats(ctypes.Structure):
    _fields_ = [("last_ms", ctypes.c_double), ("last_updates", ctypes.c_int64), ("last_launches", ctypes.c_int32),
                ("t_block", ctypes.c_int32), ("tile_x", ctypes.c_int32), ("tile_y", ctypes.c_int32),
                ("zchunk", ctypes.c_int32), ("parity", ctypes.c_int32), ("device_bytes", ctypes.c_int64),
                ("pitch_x", ctypes.c_int32), ("rows_y", ctypes.c_int32), ("planes_z", ctypes.c_int32),
                ("x_offset", ctypes.c_int32), ("tile_variant", ctypes.c_int32), ("xcd_order", ctypes.c_int32)]


class VModel(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("nlayers", ctypes.c_int32), ("nblobs", ctypes.c_int32),
                ("vmin", ctypes.c_double), ("vmax", ctypes.c_double), ("pert", ctypes.c_double),
                ("dt_over_h", ctypes.c_double), ("anomaly", ctypes.c_double),
                ("anomaly_center", ctypes.c_double * 3), ("anomaly_sigma", ctypes.c_double),
                ("round_bf16", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("last_ms", ctypes.c_double), ("last_updates", ctypes.c_int64), ("last_launches", ctypes.c_int32),
                ("nproblems", ctypes.c_int32), ("items", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("zchunk", ctypes.c_int32), ("tile_x", ctypes.c_int32),
                ("tile_y", ctypes.c_int32), ("nvariants", ctypes.c_int32), ("time_block", ctypes.c_int32),
                ("block_variant", ctypes.c_int32), ("last_mode", ctypes.c_int32)]


EXPORTS = ["gr_create", "gr_destroy", "gr_last_error", "gr_upload", "gr_download", "gr_step", "gr_run_ts",
           "gr_set_source", "gr_get_stats", "gr_mwd_round_nt", "gr_sweep_dev", "gr_field_ptr", "gr_init_fields",
           "gr_init_coef_uniform", "gr_compare_fields", "gr_compare_fields_static", "gr_set_config", "gr_set_xcd_order", "gr_set_streaming",
           "gr_set_graph", "gr_autotune",
           "gr_init_roc2_model",
           "gr_download_roc2", "gr_batch_create", "gr_batch_destroy", "gr_batch_problem", "gr_batch_set_config",
           "gr_batch_step", "gr_batch_set_time_block", "gr_batch_set_receivers", "gr_batch_autotune", "gr_batch_get_stats", "gr_shot_create", "gr_shot_destroy",
           "gr_set_roc2_from_slowness2", "gr_shot_forward", "gr_shot_adjoint", "gr_shot_take_checkpoints"]

_lib = None


def lib():
    """Load libgirih_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # Bind to the HIP runtime torch already ships when torch is present, so the
        # process holds a single HIP runtime (torch's libamdhip64 has the same soname).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgirih_gpu.so not built: run `python -m girih_amd.build` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, I = ctypes.c_void_p, ctypes.c_int
        IP = ctypes.POINTER(ctypes.c_int)
        sig = {
            "gr_create": ([ctypes.POINTER(Desc), ctypes.POINTER(P)], I),
            "gr_destroy": ([P], I),
            "gr_last_error": ([], ctypes.c_char_p),
            "gr_upload": ([P, IP, P, P, P, P], I),
            "gr_download": ([P, IP, P, P], I),
            "gr_step": ([P, I], I),
            "gr_run_ts": ([P, I, I, I, IP], I),
            "gr_set_source": ([P, I, I, I, ctypes.POINTER(ctypes.c_double), I], I),
            "gr_get_stats": ([P, ctypes.POINTER(Stats)], I),
            "gr_mwd_round_nt": ([I, I], I),
            "gr_sweep_dev": ([I, I, IP, I, I, I, I, I, I, P, P, P, P], I),
            "gr_field_ptr": ([P, I, ctypes.POINTER(P)], I),
            "gr_init_fields": ([P, I, ctypes.c_uint64], I),
            "gr_init_coef_uniform": ([P, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double], I),
            "gr_compare_fields": ([P, I, P, I, ctypes.POINTER(ctypes.c_double)], I),
            "gr_compare_fields_static": ([P, I, P, I, ctypes.POINTER(ctypes.c_double)], I),
            "gr_set_config": ([P, I, I, I], I),
            "gr_set_xcd_order": ([P, I], I),
            "gr_set_streaming": ([P, I], I),
            "gr_set_graph": ([P, I], I),
            "gr_autotune": ([P, I, ctypes.POINTER(Stats)], I),
            "gr_init_roc2_model": ([P, ctypes.POINTER(VModel)], I),
            "gr_download_roc2": ([P, IP, P], I),
            "gr_batch_create": ([ctypes.POINTER(Desc), I, ctypes.POINTER(P)], I),
            "gr_batch_destroy": ([P], I),
            "gr_batch_problem": ([P, I, ctypes.POINTER(P)], I),
            "gr_batch_set_config": ([P, I, I], I),
            "gr_batch_step": ([P, I, I], I),
            "gr_batch_set_time_block": ([P, I, I], I),
            "gr_batch_set_receivers": ([P, I, I, P, I], I),
            "gr_batch_autotune": ([P, I, I, IP, IP], I),
            "gr_batch_get_stats": ([P, ctypes.POINTER(BatchStats)], I),
            "gr_shot_create": ([P, I, I, I, ctypes.c_double, ctypes.POINTER(P)], I),
            "gr_shot_destroy": ([P], I),
            "gr_set_roc2_from_slowness2": ([P, P, ctypes.c_double], I),
            "gr_shot_forward": ([P, P], I),
            "gr_shot_adjoint": ([P, P, P], I),
            "gr_shot_take_checkpoints": ([P, P], I),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != GR_OK:
        raise GirihError(rc, lib().gr_last_error().decode())
    return rc


def mwd_round_nt(nt, t_dim):
    return lib().gr_mwd_round_nt(int(nt), int(t_dim))


def np_dtype(dtype):
    return np.float64 if dtype == GR_F64 else np.float32


def ldomain_shape(kernel, nx, ny, nz, alignment=8):
    """girih's padded local domain (utils.c:367-374) as (nnx, nny, nnz)."""
    R = RADIUS[kernel]
    nnx = nx + 2 * R
    pad = (-nnx) % alignment
    return (nnx + pad, ny + 2 * R, nz + 2 * R)


def _ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def make_desc(kernel, nx, ny, nz, dtype=GR_F64, flags=0, t_block=1, roc2=0.0, coef=None):
    d = Desc()
    d.kernel, d.dtype, d.nx, d.ny, d.nz = kernel, dtype, nx, ny, nz
    d.flags, d.t_block, d.roc2 = flags, t_block, roc2
    c = G_COEF if coef is None else coef
    for i, x in enumerate(list(c)[:11]):
        d.coef[i] = float(x)
    return d


def vmodel(seed, dt_over_h, nlayers=8, nblobs=6, vmin=1500.0, vmax=4500.0, pert=0.05, anomaly=0.0,
           anomaly_center=(0.0, 0.0, 0.0), anomaly_sigma=1.0, round_bf16=False):
    """A layered velocity model (gr_vmodel): nlayers layers along z from vmin to
    vmax m/s, random interfaces and a +-pert Gaussian-blob perturbation."""
    m = VModel()
    m.seed, m.nlayers, m.nblobs = seed, nlayers, nblobs
    m.vmin, m.vmax, m.pert, m.dt_over_h = vmin, vmax, pert, dt_over_h
    m.anomaly, m.anomaly_sigma = anomaly, anomaly_sigma
    for d in range(3):
        m.anomaly_center[d] = float(anomaly_center[d])
    m.round_bf16 = 1 if round_bf16 else 0
    return m


class Stepper:
    """One stencil problem resident in HBM (a girih time-stepper target)."""

    def __init__(self, kernel, nx, ny, nz, dtype=GR_F64, flags=0, t_block=1, roc2=0.0, coef=None, _handle=None):
        self.kernel, self.nx, self.ny, self.nz = kernel, nx, ny, nz
        self.dtype, self.flags = dtype, flags
        self.R = RADIUS.get(kernel, 1)
        self._owned = _handle is None
        if _handle is not None:  # a problem owned by a Batch
            self._h = _handle
            return
        d = make_desc(kernel, nx, ny, nz, dtype, flags, t_block, roc2, coef)
        self._h = ctypes.c_void_p()
        check(lib().gr_create(ctypes.byref(d), ctypes.byref(self._h)))

    def close(self):
        if self._h and self._owned:
            lib().gr_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def np_dtype(self):
        return np_dtype(self.dtype)

    def _ld(self, a):
        nnz, nny, nnx = a.shape
        return (ctypes.c_int * 3)(nnx, nny, nnz)

    def upload(self, U1, U2, U3=None, coef=None):
        dt = self.np_dtype
        for a in (U1, U2, U3, coef):
            if a is not None and a.dtype != dt:
                raise TypeError("expected %s arrays" % np.dtype(dt))
        check(lib().gr_upload(self._h, self._ld(U1), _ptr(U1), _ptr(U2), _ptr(U3), _ptr(coef)))

    def download(self, shape=None):
        if shape is None:
            nnx, nny, nnz = ldomain_shape(self.kernel, self.nx, self.ny, self.nz)
            shape = (nnz, nny, nnx)
        U1 = np.zeros(shape, dtype=self.np_dtype)
        U2 = np.zeros(shape, dtype=self.np_dtype)
        check(lib().gr_download(self._h, self._ld(U1), _ptr(U1), _ptr(U2)))
        return U1, U2

    def step(self, napply):
        check(lib().gr_step(self._h, int(napply)))

    def run_ts(self, target_ts, nt, t_dim=1):
        out = ctypes.c_int(0)
        check(lib().gr_run_ts(self._h, int(target_ts), int(nt), int(t_dim), ctypes.byref(out)))
        return out.value

    def set_source(self, i, j, k, values):
        vals = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().gr_set_source(self._h, i, j, k, vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  int(vals.size)))

    def init_fields(self, mode=0, seed=0):
        check(lib().gr_init_fields(self._h, int(mode), ctypes.c_uint64(seed)))

    def init_coef_uniform(self, seed, lo0, hi0, lo1, hi1):
        check(lib().gr_init_coef_uniform(self._h, ctypes.c_uint64(seed), lo0, hi0, lo1, hi1))

    def stats(self):
        s = Stats()
        check(lib().gr_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def compare(self, field, other, other_field):
        """FP64 device reduction: {sum (a-b)^2, sum b^2, max|a-b|, max|b|} over the interior."""
        out = (ctypes.c_double * 4)()
        check(lib().gr_compare_fields(self._h, int(field), other._h, int(other_field), out))
        return {"sq_diff": out[0], "sq_ref": out[1], "max_diff": out[2], "max_ref": out[3],
                "rel_l2": (out[0] / out[1]) ** 0.5 if out[1] > 0 else float("nan")}

    def compare_static(self, field, other, other_field):
        """compare() through gr_compare_fields_static (no device allocation)."""
        out = (ctypes.c_double * 4)()
        check(lib().gr_compare_fields_static(self._h, int(field), other._h, int(other_field), out))
        return {"sq_diff": out[0], "sq_ref": out[1], "max_diff": out[2], "max_ref": out[3],
                "rel_l2": (out[0] / out[1]) ** 0.5 if out[1] > 0 else float("nan")}

    def set_config(self, tile_variant=-1, zchunk=-1, t_block=-1):
        check(lib().gr_set_config(self._h, int(tile_variant), int(zchunk), int(t_block)))

    def set_xcd_order(self, order):
        """Workgroup order: 0 plain, 1 XCD-contiguous runs of the whole grid, 2 XCD bands
        of tile rows within each z chunk (gr_set_xcd_order)."""
        check(lib().gr_set_xcd_order(self._h, int(order)))

    def set_streaming(self, mode):
        """Cache policy: -1 by working-set size (default), 0 normal accesses, 1 non-temporal
        loads and stores for the data a step touches once (results do not change)."""
        check(lib().gr_set_streaming(self._h, int(mode)))

    def set_graph(self, on=True):
        """Replay gr_step's launch sequence from a HIP graph (small, launch-bound grids)."""
        check(lib().gr_set_graph(self._h, 1 if on else 0))

    def autotune(self, steps=4):
        s = Stats()
        check(lib().gr_autotune(self._h, int(steps), ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def field_ptr(self, which):
        p = ctypes.c_void_p()
        check(lib().gr_field_ptr(self._h, int(which), ctypes.byref(p)))
        return p.value

    def init_roc2_model(self, model):
        check(lib().gr_init_roc2_model(self._h, ctypes.byref(model)))

    def download_roc2(self, shape=None):
        if shape is None:
            nnx, nny, nnz = ldomain_shape(self.kernel, self.nx, self.ny, self.nz)
            shape = (nnz, nny, nnx)
        U3 = np.zeros(shape, dtype=self.np_dtype)
        check(lib().gr_download_roc2(self._h, self._ld(U3), _ptr(U3)))
        return U3


class Batch:
    """Independent kernel-0 problems of uneven size advanced together, one
    persistent work-queue launch per time step or per two-step block (gr_batch_*)."""

    def __init__(self, sizes, flags=GR_FAST | GR_COEF_BF16, coef=None):
        self.sizes = [tuple(int(x) for x in s) for s in sizes]
        descs = (Desc * len(self.sizes))()
        for i, (nx, ny, nz) in enumerate(self.sizes):
            descs[i] = make_desc(0, nx, ny, nz, GR_F32, flags, 1, 0.0, coef)
        self.flags = flags
        self._b = ctypes.c_void_p()
        check(lib().gr_batch_create(descs, len(self.sizes), ctypes.byref(self._b)))
        self.problems = []
        for i, (nx, ny, nz) in enumerate(self.sizes):
            h = ctypes.c_void_p()
            check(lib().gr_batch_problem(self._b, i, ctypes.byref(h)))
            self.problems.append(Stepper(0, nx, ny, nz, GR_F32, flags, _handle=h))

    def close(self):
        if self._b:
            for p in self.problems:
                p.close()
            lib().gr_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.problems)

    def __getitem__(self, i):
        return self.problems[i]

    def step(self, napply, use_graph=False, mode=None):
        """mode 0: a launch per step, 1: replayed from a graph (use_graph=True),
        2: one cooperative launch for the whole call (gr_batch_step)."""
        m = (1 if use_graph else 0) if mode is None else int(mode)
        check(lib().gr_batch_step(self._b, int(napply), m))

    def set_config(self, variant=-1, zchunk=-1):
        check(lib().gr_batch_set_config(self._b, int(variant), int(zchunk)))

    def set_time_block(self, tblock, block_variant=-1):
        """1: one launch per step; 2: two-step blocks, block_variant 0 (registers
        uncapped) or 1 (capped at 128) (gr_batch_set_time_block)."""
        check(lib().gr_batch_set_time_block(self._b, int(tblock), int(block_variant)))

    def set_receivers(self, i, rec_k, trace_ptr, max_steps):
        """Record plane rec_k of problem i after every application into the device
        array at trace_ptr, FP32 [max_steps][ny][nx] (gr_batch_set_receivers); a torch
        tensor is taken by its data pointer, None removes the receiver."""
        ptr = trace_ptr.data_ptr() if hasattr(trace_ptr, "data_ptr") else trace_ptr
        check(lib().gr_batch_set_receivers(self._b, int(i), int(rec_k), ctypes.c_void_p(ptr) if ptr else None,
                                           int(max_steps)))

    def autotune(self, steps=4, use_graph=False, mode=None):
        v, z = ctypes.c_int(0), ctypes.c_int(0)
        m = (1 if use_graph else 0) if mode is None else int(mode)
        check(lib().gr_batch_autotune(self._b, int(steps), m, ctypes.byref(v), ctypes.byref(z)))
        return v.value, z.value

    def stats(self):
        s = BatchStats()
        check(lib().gr_batch_get_stats(self._b, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in BatchStats._fields_}


def sweep_dev(kernel, dtype, shape_xyz, box, coef, u_ptr, v_ptr, roc2_ptr=None):
    """Level-2 entry: one sweep on device arrays in girih layout (KERNEL_SIG).

    coef: numpy constants (kernels 0, 1) or a device pointer (kernels 2-5)."""
    shape = (ctypes.c_int * 3)(*shape_xyz)
    xb, yb, zb, xe, ye, ze = box
    if isinstance(coef, np.ndarray):
        coef = np.ascontiguousarray(coef, dtype=np_dtype(dtype))
        cptr = _ptr(coef)
    else:
        cptr = ctypes.c_void_p(coef)
    check(lib().gr_sweep_dev(kernel, dtype, shape, xb, yb, zb, xe, ye, ze, cptr, ctypes.c_void_p(u_ptr),
                             ctypes.c_void_p(v_ptr), ctypes.c_void_p(roc2_ptr) if roc2_ptr else None))
