// gr_adjoint.hip -- seismic shots on the kernel-0 wave stepper: receiver traces,
// and the gradient of any function of the traces with respect to the squared
// slowness m = 1/v^2, by the discrete adjoint with checkpointing.
//
// Forward (girih's kernel 0 with a per-cell roc2 = (v dt/h)^2 = dt^2/(h^2 m),
// src/kernels/stencils.c:27-41, source as nb_naive_ts.c:190,196):
//   x_0 = x_1 = 0,  x_{a+2} = 2 x_{a+1} - x_a + r (L x_{a+1}) + s_a e_src,
//   trace a = P x_{a+2} (the interior of plane rec_k),       a = 0 .. nt-1.
// L is the symmetric star stencil on the interior (girih's halo is a zero,
// never-updated boundary).  For a cotangent res_a of trace a, the adjoint in the
// scaled variable nu = r * mu (mu the Lagrange multiplier of x) is the SAME
// leapfrog, run backwards with the receivers as sources:
//   nu_{a+2} = 2 nu_{a+3} - nu_{a+4} + r (L nu_{a+3}) + r P^T res_a,  nu_{nt+2} = nu_{nt+3} = 0,
// so the forward sweep kernel steps it unchanged.  With r L x_{a+1} =
// x_{a+2} - 2x_{a+1} + x_a - s_a e_src and dr/dm = -r^2 h^2/dt^2:
//   dJ/dm = -(h/dt)^2 * sum_a nu_{a+2} * ((x_{a+2} - x_{a+1}) - (x_{a+1} - x_a) - s_a e_src),
// a pointwise accumulation (differences in FP64, sum in the field type) with no
// second stencil pass.
//
// The forward pass stores (x_{cK}, x_{cK+1}) every K applications; the backward
// pass recomputes each segment's K states from its checkpoint into a buffer and
// walks it backwards.  Memory: 2*ceil(nt/K) + 2 fields for the forward pass;
// the first backward pass adds 2K + 4 fields (the segment's states and adjoint
// fields) and a dense accumulator, which takes a segment's gradient terms in one pass.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <vector>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"
#include "gr_sweep.hpp"

namespace {

using gr::Problem;

template <typename T>
__global__ void record_kernel(const T *x, long long sy, long long plane_off, int nx, int ny, int R, T *rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < nx) rec[(long long)j * nx + i] = x[plane_off + (long long)(j + R) * sy + (i + R)];
}

// nu(plane rec_k) += roc2 * res  (the adjoint source of the traces)
template <typename T, typename C>
__global__ void inject_kernel(T *nu, const C *roc2, long long sy, long long plane_off, int nx, int ny, int R,
                              const T *res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const long long o = plane_off + (long long)(j + R) * sy + (i + R);
  nu[o] = nu[o] + (T)(float)roc2[o] * res[(long long)j * nx + i];
}
template <typename T>
__global__ void inject_kernel_same(T *nu, const T *roc2, long long sy, long long plane_off, int nx, int ny, int R,
                                   const T *res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const long long o = plane_off + (long long)(j + R) * sy + (i + R);
  nu[o] = nu[o] + roc2[o] * res[(long long)j * nx + i];
}

// g += nu_{a+2} * ((x_{a+2} - x_{a+1}) - (x_{a+1} - x_a) - s_a e_src) over the
// interior (g dense [nz][ny][nx]) for the steps a = a1-1 down to a0 of one checkpoint
// segment, in one pass: per cell the terms are added in the order of the backward
// walk, as the CPU adjoint adds them, but g is read and written once per segment and
// each state and adjoint field once.  The differences are formed in FP64 (exact for
// FP32 states); g has the field type: an FP32 shot's gradient error is set by its FP32
// traces, far above FP32 accumulation over the steps.  States x_{a0} .. x_{a1+1} are
// fields xs + t * fstride; nu_{a+2} is field nus + (a - a0) * fstride.  The source
// term applies at cell src_idx for steps a < nsrc.
template <typename T>
__global__ void grad_segment_kernel(T *g, const T *__restrict__ nus, const T *__restrict__ xs, long long fstride,
                                    long long sy, long long sz, int nx, int ny, int R, int nsteps, const T *src,
                                    long long src_idx, int a0, int nsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const long long o = (long long)(k + R) * sz + (long long)(j + R) * sy + (i + R);
  T &gc = g[((long long)k * ny + j) * nx + i];
  T acc = gc;
  const bool at_src = src && o == src_idx;
  double x2 = (double)xs[(long long)(nsteps + 1) * fstride + o];  // x_{a+2} for a = a1 - 1
  double x1 = (double)xs[(long long)nsteps * fstride + o];
  // the loads of QB steps are issued together (each cell's terms are still added one
  // step at a time, in the walk's order)
  constexpr int QB = 8;
  for (int t0 = nsteps - 1; t0 >= 0; t0 -= QB) {  // t = a - a0
    T xq[QB], nq[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q)
      if (t0 - q >= 0) {
        xq[q] = __builtin_nontemporal_load(xs + (long long)(t0 - q) * fstride + o);
        nq[q] = __builtin_nontemporal_load(nus + (long long)(t0 - q) * fstride + o);
      }
#pragma unroll
    for (int q = 0; q < QB; ++q)
      if (t0 - q >= 0) {
        const int t = t0 - q;
        const double x0 = (double)xq[q];
        double d = (x2 - x1) - (x1 - x0);
        if (at_src && a0 + t < nsrc) d = d - (double)src[a0 + t];
        acc = (T)((double)acc + (double)nq[q] * d);
        x2 = x1;
        x1 = x0;
      }
  }
  gc = acc;
}

template <typename T>
__global__ void grad_final_kernel(const T *g, T *out, long long n, double scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (T)(scale * (double)g[t]);
}

// dst <- src cell by cell over the whole padded grid (both fields of one grid size;
// pitches differ with the element size), rounding to nearest when D is narrower
template <typename S, typename D>
__global__ void convert_field_kernel(D *dst, const S *src, int nnx, long long spitch, long long dpitch,
                                     long long total) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long long row = t / nnx;
  const int i = (int)(t - row * nnx);
  dst[row * dpitch + i] = (D)src[row * spitch + i];
}

template <typename T>
__global__ void zero_kernel(T *g, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) g[t] = T(0);
}

// roc2 = dt_over_h^2 / m over the whole field (halo cells take the nearest interior m)
template <typename T, typename C>
__global__ void roc2_from_m_kernel(C *roc2, const T *m, int nx, int ny, int nz, int R, long long pitch,
                                   long long rows, long long total, double k2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int nnx = nx + 2 * R;
  const long long row = t / nnx;
  const int i = (int)(t - row * nnx);
  const int j = (int)(row % rows), k = (int)(row / rows);
  const int ii = min(max(i - R, 0), nx - 1), jj = min(max(j - R, 0), ny - 1), kk = min(max(k - R, 0), nz - 1);
  const double mv = (double)m[((long long)kk * ny + jj) * nx + ii];
  const double r = k2 / mv;
  if constexpr (std::is_same<C, __hip_bfloat16>::value)
    roc2[row * pitch + i] = __float2bfloat16((float)r);
  else
    roc2[row * pitch + i] = (C)r;
}

}  // namespace

struct gr_shot {
  gr_handle *h = nullptr;
  int rec_k = 0, nt = 0, ckpt = 0, nck = 0;
  double dt_over_h = 0;
  std::vector<void *> allocs;  // everything below lives in these
  std::vector<void *> ck;      // 2*nck checkpoint fields (x_{cK}, x_{cK+1})
  void *seg = nullptr;         // ckpt+2 recompute fields x_{a0} .. x_{a1+1}, contiguous
  void *nus = nullptr;         // ckpt+2 adjoint fields nu_{a0+2} .. nu_{a1+3}, contiguous
  void *x[2] = {nullptr, nullptr};  // forward ping-pong
  void *g = nullptr;  // gradient accumulator, field type, dense interior
  bool have_forward = false;
};

namespace {

int shot_fail(hipError_t e, const char *what) { return gr::set_errorf(GR_EHIP, "%s: %s", what, hipGetErrorString(e)); }

// field base (x offset applied) of a raw field allocation
void *fbase(const Problem &P, void *alloc) { return (char *)alloc + (size_t)P.xoff * gr::elem_size(P.dtype); }

template <typename T>
hipError_t record(const Problem &P, const void *x, int rec_k, void *rec, hipStream_t st) {
  const long long sz = (long long)P.pitch * P.rows;
  hipLaunchKernelGGL(record_kernel<T>, dim3((P.nx + 255) / 256, P.ny), dim3(256), 0, st, (const T *)x,
                     (long long)P.pitch, (long long)rec_k * sz, P.nx, P.ny, P.R, (T *)rec);
  return hipGetLastError();
}

hipError_t record_any(const Problem &P, const void *x, int rec_k, void *rec, hipStream_t st) {
  return P.dtype == GR_F64 ? record<double>(P, x, rec_k, rec, st) : record<float>(P, x, rec_k, rec, st);
}

hipError_t inject_any(const Problem &P, void *nu, int rec_k, const void *res, hipStream_t st) {
  const long long sz = (long long)P.pitch * P.rows;
  const dim3 grid((P.nx + 255) / 256, P.ny);
  if (P.dtype == GR_F64)
    hipLaunchKernelGGL(inject_kernel_same<double>, grid, dim3(256), 0, st, (double *)nu, (const double *)P.roc2,
                       (long long)P.pitch, (long long)rec_k * sz, P.nx, P.ny, P.R, (const double *)res);
  else if (P.flags & GR_COEF_BF16)
    hipLaunchKernelGGL((inject_kernel<float, __hip_bfloat16>), grid, dim3(256), 0, st, (float *)nu,
                       (const __hip_bfloat16 *)P.roc2, (long long)P.pitch, (long long)rec_k * sz, P.nx, P.ny, P.R,
                       (const float *)res);
  else
    hipLaunchKernelGGL(inject_kernel_same<float>, grid, dim3(256), 0, st, (float *)nu, (const float *)P.roc2,
                       (long long)P.pitch, (long long)rec_k * sz, P.nx, P.ny, P.R, (const float *)res);
  return hipGetLastError();
}

// the accumulation of one segment's nsteps steps starting at a0 (grad_segment_kernel)
hipError_t grad_segment(gr_handle *h, void *g, const void *nus, const void *xs, int nsteps, int a0, hipStream_t st) {
  const Problem &P = gr_internal_problem(h);
  const long long sz = (long long)P.pitch * P.rows;
  const dim3 grid((P.nx + 255) / 256, P.ny, P.nz);
  int nsrc = 0;
  long long sidx = -1;
  const void *src = gr_internal_source(h, &nsrc, &sidx);
  if (!src || sidx < 0) {
    src = nullptr;
    sidx = -1;
  }
  if (P.dtype == GR_F64)
    hipLaunchKernelGGL(grad_segment_kernel<double>, grid, dim3(256), 0, st, (double *)g, (const double *)nus,
                       (const double *)xs, (long long)P.field_elems, (long long)P.pitch, sz, P.nx, P.ny, P.R, nsteps,
                       (const double *)src, sidx, a0, nsrc);
  else
    hipLaunchKernelGGL(grad_segment_kernel<float>, grid, dim3(256), 0, st, (float *)g, (const float *)nus,
                       (const float *)xs, (long long)P.field_elems, (long long)P.pitch, sz, P.nx, P.ny, P.R, nsteps,
                       (const float *)src, sidx, a0, nsrc);
  return hipGetLastError();
}

// source value pointer of application a (null when none)
const void *src_of(gr_handle *h, long long a, long long *sidx) {
  int n = 0;
  const void *s = gr_internal_source(h, &n, sidx);
  if (!s || *sidx < 0 || a >= n) {
    *sidx = -1;
    return nullptr;
  }
  return (const char *)s + (size_t)a * gr::elem_size(gr_internal_problem(h).dtype);
}

}  // namespace

extern "C" int gr_shot_destroy(gr_shot *s) {
  if (!s) return GR_OK;
  if (s->h) (void)hipStreamSynchronize(gr_internal_stream(s->h));
  for (void *p : s->allocs)
    if (p) (void)hipFree(p);
  delete s;
  return GR_OK;
}

extern "C" int gr_shot_create(gr_handle *h, int rec_k, int nt, int ckpt, double dt_over_h, gr_shot **out) {
  if (!h || !out) return gr::set_error(GR_EINVAL, "gr_shot_create: null argument");
  *out = nullptr;
  const Problem &P = gr_internal_problem(h);
  if (P.kernel != 0 || !P.roc2)
    return gr::set_error(GR_EINVAL, "gr_shot_create: shots run kernel 0 with a roc2 field (not GR_ROC2_SCALAR)");
  if (rec_k < P.R || rec_k >= P.nz + P.R)
    return gr::set_errorf(GR_EINVAL, "gr_shot_create: receiver plane %d is not an interior plane", rec_k);
  if (nt < 1 || ckpt < 1 || !(dt_over_h > 0))
    return gr::set_error(GR_EINVAL, "gr_shot_create: need nt >= 1, ckpt >= 1, dt_over_h > 0");
  gr_shot *s = new gr_shot();
  s->h = h;
  s->rec_k = rec_k;
  s->nt = nt;
  s->ckpt = std::min(ckpt, nt);
  s->nck = (nt + s->ckpt - 1) / s->ckpt;
  s->dt_over_h = dt_over_h;
  const size_t fbytes = (size_t)P.field_elems * gr::elem_size(P.dtype);
  hipStream_t st = gr_internal_stream(h);
  auto field = [&](void **p) -> hipError_t {
    hipError_t e = hipMalloc(p, fbytes);
    if (e == hipSuccess) s->allocs.push_back(*p);
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, fbytes, st);  // girih's zero halo, never written
    return e;
  };
  hipError_t e = hipSuccess;
  for (int i = 0; i < 2 * s->nck && e == hipSuccess; ++i) {
    void *p = nullptr;
    if ((e = field(&p)) == hipSuccess) s->ck.push_back(p);
  }
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = field(&s->x[i]);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    gr_shot_destroy(s);
    return gr::set_errorf(GR_ENOMEM, "gr_shot_create: device allocation failed: %s", hipGetErrorString(e));
  }
  *out = s;
  return GR_OK;
}

extern "C" int gr_set_roc2_from_slowness2(gr_handle *h, const void *m_dev, double dt_over_h) {
  if (!h || !m_dev) return gr::set_error(GR_EINVAL, "gr_set_roc2_from_slowness2: null argument");
  const Problem &P = gr_internal_problem(h);
  if (P.kernel != 0 || !P.roc2) return gr::set_error(GR_EINVAL, "gr_set_roc2_from_slowness2: no roc2 field");
  if (!(dt_over_h > 0)) return gr::set_error(GR_EINVAL, "gr_set_roc2_from_slowness2: dt_over_h must be > 0");
  const long long total = (long long)(P.nx + 2 * P.R) * P.rows * P.planes;
  const unsigned nb = (unsigned)((total + 255) / 256);
  const double k2 = dt_over_h * dt_over_h;
  hipStream_t st = gr_internal_stream(h);
  if (P.dtype == GR_F64)
    hipLaunchKernelGGL((roc2_from_m_kernel<double, double>), dim3(nb), dim3(256), 0, st, (double *)P.roc2,
                       (const double *)m_dev, P.nx, P.ny, P.nz, P.R, (long long)P.pitch, (long long)P.rows, total, k2);
  else if (P.flags & GR_COEF_BF16)
    hipLaunchKernelGGL((roc2_from_m_kernel<float, __hip_bfloat16>), dim3(nb), dim3(256), 0, st,
                       (__hip_bfloat16 *)P.roc2, (const float *)m_dev, P.nx, P.ny, P.nz, P.R, (long long)P.pitch,
                       (long long)P.rows, total, k2);
  else
    hipLaunchKernelGGL((roc2_from_m_kernel<float, float>), dim3(nb), dim3(256), 0, st, (float *)P.roc2,
                       (const float *)m_dev, P.nx, P.ny, P.nz, P.R, (long long)P.pitch, (long long)P.rows, total, k2);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return shot_fail(e, "gr_set_roc2_from_slowness2");
  return GR_OK;
}

extern "C" int gr_shot_forward(gr_shot *s, void *rec_dev) {
  if (!s || !rec_dev) return gr::set_error(GR_EINVAL, "gr_shot_forward: null argument");
  gr_handle *h = s->h;
  const Problem &P = gr_internal_problem(h);
  const gr::LaunchCfg &lc = gr_internal_launch_cfg(h);
  hipStream_t st = gr_internal_stream(h);
  const size_t fbytes = (size_t)P.field_elems * gr::elem_size(P.dtype);
  const size_t plane = (size_t)P.nx * P.ny * gr::elem_size(P.dtype);
  hipError_t e = hipSuccess;
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipMemsetAsync(s->x[i], 0, fbytes, st);
  void *prev = s->x[0], *cur = s->x[1];  // x_a, x_{a+1}
  for (int a = 0; a < s->nt && e == hipSuccess; ++a) {
    if (a % s->ckpt == 0) {
      const int c = a / s->ckpt;
      e = hipMemcpyAsync(s->ck[2 * c], prev, fbytes, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(s->ck[2 * c + 1], cur, fbytes, hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) break;
    }
    long long sidx = -1;
    const void *sv = src_of(h, a, &sidx);
    e = gr::sweep_fields(P, fbase(P, prev), fbase(P, cur), sv, sidx, lc, st);  // prev <- x_{a+2}
    if (e == hipSuccess) e = record_any(P, fbase(P, prev), s->rec_k, (char *)rec_dev + (size_t)a * plane, st);
    std::swap(prev, cur);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return shot_fail(e, "gr_shot_forward");
  s->have_forward = true;
  return GR_OK;
}

extern "C" int gr_shot_take_checkpoints(gr_shot *dst, const gr_shot *src) {
  if (!dst || !src) return gr::set_error(GR_EINVAL, "gr_shot_take_checkpoints: null argument");
  if (!src->have_forward) return gr::set_error(GR_EINVAL, "gr_shot_take_checkpoints: the source shot has no forward run");
  const Problem &D = gr_internal_problem(dst->h), &S = gr_internal_problem(src->h);
  if (D.nx != S.nx || D.ny != S.ny || D.nz != S.nz || dst->nt != src->nt || dst->ckpt != src->ckpt ||
      dst->rec_k != src->rec_k)
    return gr::set_error(GR_EINVAL, "gr_shot_take_checkpoints: the shots differ in grid, nt, ckpt or receiver plane");
  if (S.dtype == GR_F32 && D.dtype == GR_F64)
    return gr::set_error(GR_EUNSUPPORTED, "gr_shot_take_checkpoints: FP32 checkpoints cannot seed an FP64 shot");
  // the source shot's stream has finished its forward run (gr_shot_forward synchronises)
  hipStream_t st = gr_internal_stream(dst->h);
  const int nnx = D.nx + 2 * D.R;
  const long long total = (long long)nnx * D.rows * D.planes;
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < dst->ck.size() && e == hipSuccess; ++i) {
    void *d = fbase(D, dst->ck[i]);
    const void *sp = fbase(S, src->ck[i]);
    if (S.dtype == GR_F64 && D.dtype == GR_F64)
      hipLaunchKernelGGL((convert_field_kernel<double, double>), dim3(nb), dim3(256), 0, st, (double *)d,
                         (const double *)sp, nnx, (long long)S.pitch, (long long)D.pitch, total);
    else if (S.dtype == GR_F64)
      hipLaunchKernelGGL((convert_field_kernel<double, float>), dim3(nb), dim3(256), 0, st, (float *)d,
                         (const double *)sp, nnx, (long long)S.pitch, (long long)D.pitch, total);
    else
      hipLaunchKernelGGL((convert_field_kernel<float, float>), dim3(nb), dim3(256), 0, st, (float *)d,
                         (const float *)sp, nnx, (long long)S.pitch, (long long)D.pitch, total);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return shot_fail(e, "gr_shot_take_checkpoints");
  dst->have_forward = true;
  return GR_OK;
}

extern "C" int gr_shot_adjoint(gr_shot *s, const void *res_dev, void *grad_dev) {
  if (!s || !res_dev || !grad_dev) return gr::set_error(GR_EINVAL, "gr_shot_adjoint: null argument");
  if (!s->have_forward) return gr::set_error(GR_EINVAL, "gr_shot_adjoint: run gr_shot_forward first");
  gr_handle *h = s->h;
  const Problem &P = gr_internal_problem(h);
  const gr::LaunchCfg &lc = gr_internal_launch_cfg(h);
  hipStream_t st = gr_internal_stream(h);
  const size_t fbytes = (size_t)P.field_elems * gr::elem_size(P.dtype);
  const size_t plane = (size_t)P.nx * P.ny * gr::elem_size(P.dtype);
  const long long ncell = (long long)P.nx * P.ny * P.nz;
  hipError_t e = hipSuccess;
  const int nslot = s->ckpt + 2;
  if (!s->seg) {  // backward-pass buffers, allocated on first use (forward-only shots never hold them)
    const size_t before = s->allocs.size();
    auto fields = [&](void **p, int n) -> hipError_t {
      hipError_t r = hipMalloc(p, fbytes * (size_t)n);
      if (r == hipSuccess) s->allocs.push_back(*p);
      if (r == hipSuccess) r = hipMemsetAsync(*p, 0, fbytes * (size_t)n, st);  // zero halo, ordered before first use
      return r;
    };
    e = fields(&s->seg, nslot);
    if (e == hipSuccess) e = fields(&s->nus, nslot);
    if (e == hipSuccess) {
      e = hipMalloc(&s->g, gr::elem_size(P.dtype) * (size_t)ncell);
      if (e == hipSuccess) s->allocs.push_back(s->g);
    }
    if (e != hipSuccess) {  // release the partial set so a later call starts clean
      for (size_t i = before; i < s->allocs.size(); ++i) (void)hipFree(s->allocs[i]);
      s->allocs.resize(before);
      s->seg = s->nus = nullptr;
      s->g = nullptr;
      return gr::set_errorf(GR_ENOMEM, "gr_shot_adjoint: device allocation failed: %s", hipGetErrorString(e));
    }
  }
  auto segf = [&](int t) { return (void *)((char *)s->seg + (size_t)t * fbytes); };  // x_{a0+t}
  auto nuf = [&](int t) { return (void *)((char *)s->nus + (size_t)t * fbytes); };   // nu_{a0+2+t}
  if (e == hipSuccess) {
    if (P.dtype == GR_F64)
      hipLaunchKernelGGL(zero_kernel<double>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, (double *)s->g,
                         ncell);
    else
      hipLaunchKernelGGL(zero_kernel<float>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, (float *)s->g,
                         ncell);
    e = hipGetLastError();
  }
  for (int c = s->nck - 1; c >= 0 && e == hipSuccess; --c) {
    const int a0 = c * s->ckpt, a1 = std::min(a0 + s->ckpt, s->nt), K = a1 - a0;
    // recompute x_{a0} .. x_{a1+1} into seg slots 0 .. K+1
    e = hipMemcpyAsync(segf(0), s->ck[2 * c], fbytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(segf(1), s->ck[2 * c + 1], fbytes, hipMemcpyDeviceToDevice, st);
    for (int a = a0; a < a1 && e == hipSuccess; ++a) {
      const int t = a - a0 + 2;
      long long sidx = -1;
      const void *sv = src_of(h, a, &sidx);
      e = gr::sweep_fields(P, fbase(P, segf(t)), fbase(P, segf(t - 1)), sv, sidx, lc, st, fbase(P, segf(t - 2)));
    }
    // nu_{a1+2}, nu_{a1+3} into slots K, K+1: zero after the last step, else the two
    // lowest of the later segment (its slots 0 and 1)
    if (e == hipSuccess) {
      if (c == s->nck - 1) {
        e = hipMemsetAsync(nuf(K), 0, fbytes, st);
        if (e == hipSuccess) e = hipMemsetAsync(nuf(K + 1), 0, fbytes, st);
      } else {  // slot 1 first: with K = 1 slot 1 is also a destination
        e = hipMemcpyAsync(nuf(K + 1), nuf(1), fbytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(nuf(K), nuf(0), fbytes, hipMemcpyDeviceToDevice, st);
      }
    }
    for (int a = a1 - 1; a >= a0 && e == hipSuccess; --a) {  // nu_{a+2} <- f(nu_{a+3}, nu_{a+4})
      const int t = a - a0;
      e = gr::sweep_fields(P, fbase(P, nuf(t)), fbase(P, nuf(t + 1)), nullptr, -1, lc, st, fbase(P, nuf(t + 2)));
      if (e == hipSuccess) e = inject_any(P, fbase(P, nuf(t)), s->rec_k, (const char *)res_dev + (size_t)a * plane, st);
    }
    if (e == hipSuccess) e = grad_segment(h, s->g, fbase(P, nuf(0)), fbase(P, segf(0)), K, a0, st);
  }
  if (e == hipSuccess) {
    const double scale = -1.0 / (s->dt_over_h * s->dt_over_h);
    const unsigned nb = (unsigned)((ncell + 255) / 256);
    if (P.dtype == GR_F64)
      hipLaunchKernelGGL(grad_final_kernel<double>, dim3(nb), dim3(256), 0, st, (const double *)s->g,
                         (double *)grad_dev, ncell, scale);
    else
      hipLaunchKernelGGL(grad_final_kernel<float>, dim3(nb), dim3(256), 0, st, (const float *)s->g, (float *)grad_dev,
                         ncell, scale);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return shot_fail(e, "gr_shot_adjoint");
  return GR_OK;
}
