"""Seismic shots on the 25-point wave stepper, differentiable in torch.

    shot = Shot(nx, ny, nz, nt, dt_over_h, src_point, src, rec_k, ckpt=50)
    traces = shot_traces(m, shot)                  # torch.autograd.Function
    loss = 0.5 * ((traces - d_obs) ** 2).sum()
    loss.backward()                                # m.grad = dJ/dm by the adjoint

m is the squared slowness 1/v^2 per interior cell, a dense CUDA tensor
[nz, ny, nx] of the field type; traces are [nt, ny, nx] (the interior of plane
rec_k after every application).  The forward and the vector-Jacobian product run
in libgirih_gpu.so (gr_shot_*): the same HIP sweep kernels as the stepper,
checkpoints every `ckpt` applications, recomputation in the backward pass.  There
is no CPU path: without the library every call raises.

MixedShot returns FP64 traces: the forward run is FP64, so a misfit residual that
is a small fraction of the traces keeps its accuracy, and the backward pass
(recomputation from the FP64 checkpoints rounded to FP32, the adjoint and the
gradient) runs in FP32.  m may be FP64 (the forward run uses it as it is) or FP32.
"""
import ctypes

import torch

from . import gpu, physics


class Shot:
    """A kernel-0 problem with a point source and a receiver plane (gr_shot)."""

    def __init__(self, nx, ny, nz, nt, dt_over_h, src_point, src, rec_k, ckpt=50, dtype=gpu.GR_F32, coef=None,
                 tune=True):
        flags = gpu.GR_FAST if dtype == gpu.GR_F32 else 0
        self.nx, self.ny, self.nz, self.nt = nx, ny, nz, nt
        self.dt_over_h = float(dt_over_h)
        self.dtype = dtype
        self.torch_dtype = torch.float32 if dtype == gpu.GR_F32 else torch.float64
        self.stepper = gpu.Stepper(0, nx, ny, nz, dtype=dtype, flags=flags,
                                   coef=physics.fd8_coef() if coef is None else coef)
        self.stepper.set_source(*src_point, src)
        if tune:  # tile shape and z chunk of the sweeps (times scratch fields; results do not depend on it)
            self.stepper.autotune(4)
            self.stepper.set_config(t_block=1)  # shots step one application at a time
        else:
            self.stepper.set_config(zchunk=64)
        self._s = ctypes.c_void_p()
        gpu.check(gpu.lib().gr_shot_create(self.stepper._h, int(rec_k), int(nt), int(ckpt), self.dt_over_h,
                                           ctypes.byref(self._s)))
        self._m_version = None  # the m of the last forward run (checkpoints belong to it)

    def close(self):
        if self._s:
            gpu.lib().gr_shot_destroy(self._s)
            self._s = ctypes.c_void_p()
        self.stepper.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t, shape):
        if not (t.is_cuda and t.dtype == self.torch_dtype and t.is_contiguous() and tuple(t.shape) == shape):
            raise ValueError("expected a contiguous CUDA %s tensor of shape %s" % (self.torch_dtype, shape))

    def forward(self, m):
        """Traces [nt, ny, nx] for squared slowness m [nz, ny, nx]."""
        self._check(m, (self.nz, self.ny, self.nx))
        traces = torch.empty((self.nt, self.ny, self.nx), dtype=self.torch_dtype, device=m.device)
        torch.cuda.current_stream(m.device).synchronize()  # m is written by torch's stream
        gpu.check(gpu.lib().gr_set_roc2_from_slowness2(self.stepper._h, ctypes.c_void_p(m.data_ptr()),
                                                       self.dt_over_h))
        gpu.check(gpu.lib().gr_shot_forward(self._s, ctypes.c_void_p(traces.data_ptr())))
        self._m_version = (m.data_ptr(), m._version)
        return traces

    def vjp(self, m, res):
        """sum_a <res[a], d traces[a] / dm>: the adjoint gradient [nz, ny, nx]."""
        res = res.contiguous()
        self._check(res, (self.nt, self.ny, self.nx))
        if self._m_version != (m.data_ptr(), m._version):  # checkpoints are of another m: recompute them
            self.forward(m)
        grad = torch.empty((self.nz, self.ny, self.nx), dtype=self.torch_dtype, device=res.device)
        torch.cuda.current_stream(res.device).synchronize()
        gpu.check(gpu.lib().gr_shot_adjoint(self._s, ctypes.c_void_p(res.data_ptr()), ctypes.c_void_p(grad.data_ptr())))
        return grad


class MixedShot:
    """FP64 forward run, FP32 backward pass (gr_shot_take_checkpoints).

    forward(m) takes the squared slowness [nz, ny, nx] as FP64 or FP32 and returns
    FP64 traces [nt, ny, nx]; vjp(m, res) takes FP64 (or FP32) cotangents and returns
    the gradient, computed in FP32, in m's dtype.  The FP32 shot's roc2 comes from m
    rounded to FP32 and its source values are rounded to FP32."""

    def __init__(self, nx, ny, nz, nt, dt_over_h, src_point, src, rec_k, ckpt=50, coef=None, tune=True):
        import numpy as np
        src = np.asarray(src, dtype=np.float64)
        self.nx, self.ny, self.nz, self.nt = nx, ny, nz, nt
        self.dt_over_h = float(dt_over_h)
        self.s64 = Shot(nx, ny, nz, nt, dt_over_h, src_point, src, rec_k, ckpt=ckpt, dtype=gpu.GR_F64, coef=coef,
                        tune=tune)
        self.s32 = Shot(nx, ny, nz, nt, dt_over_h, src_point, src.astype(np.float32).astype(np.float64), rec_k,
                        ckpt=ckpt, dtype=gpu.GR_F32, coef=coef, tune=tune)
        self._m_version = None

    def close(self):
        self.s64.close()
        self.s32.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, m):
        """FP64 traces for the squared slowness m (FP64, or FP32 widened exactly)."""
        if m.dtype not in (torch.float32, torch.float64):
            raise ValueError("expected an FP32 or FP64 tensor")
        m64 = m.double()  # m itself when it is FP64
        self.s64._check(m64, (self.nz, self.ny, self.nx))
        traces = self.s64.forward(m64)
        m32 = m.float().contiguous()
        torch.cuda.current_stream(m.device).synchronize()
        gpu.check(gpu.lib().gr_set_roc2_from_slowness2(self.s32.stepper._h, ctypes.c_void_p(m32.data_ptr()),
                                                       self.dt_over_h))
        gpu.check(gpu.lib().gr_shot_take_checkpoints(self.s32._s, self.s64._s))
        self._m_version = (m.data_ptr(), m._version, m.dtype)
        return traces

    def vjp(self, m, res):
        """Gradient [nz, ny, nx] in m's dtype (computed in FP32) for cotangents res [nt, ny, nx]."""
        if self._m_version != (m.data_ptr(), m._version, m.dtype):
            self.forward(m)
        res = res.to(torch.float32).contiguous()
        self.s32._check(res, (self.nt, self.ny, self.nx))
        grad = torch.empty((self.nz, self.ny, self.nx), dtype=torch.float32, device=res.device)
        torch.cuda.current_stream(res.device).synchronize()
        gpu.check(gpu.lib().gr_shot_adjoint(self.s32._s, ctypes.c_void_p(res.data_ptr()),
                                            ctypes.c_void_p(grad.data_ptr())))
        return grad.to(m.dtype)


class _ShotTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, shot):
        ctx.shot = shot
        ctx.save_for_backward(m)
        return shot.forward(m.detach())

    @staticmethod
    def backward(ctx, res):
        (m,) = ctx.saved_tensors
        return ctx.shot.vjp(m.detach(), res), None


def shot_traces(m, shot):
    """Differentiable receiver traces of `shot` (a Shot or a MixedShot) for squared
    slowness m."""
    return _ShotTraces.apply(m, shot)
