// gr_model.hip -- device-side velocity models for kernel 0 (the roc2 = (v*dt/h)^2
// field that girih keeps in U3, src/utils.c:159-162) and the copy of roc2 back to
// the host.
//
// A layered model: `nlayers` horizontal layers stacked along z with speeds rising
// linearly from vmin to vmax, interface depths drawn from the seed; the speed is
// scaled by 1 + pert * clamp(sum of `nblobs` signed Gaussian blobs, -1, 1) and by
// an optional centred Gaussian anomaly 1 + anomaly * exp(-|x - c|^2 / (2 sigma^2)).
// Every random parameter is drawn on the host from the seed (splitmix64), so a
// model is a pure function of (seed, grid, parameters) whatever the field type.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"

namespace {

constexpr int kMaxLayers = 16, kMaxBlobs = 16;

struct ModelParams {
  int nlayers, nblobs, round_bf16;
  int nx, ny, nz, R;
  long long pitch, rows;
  double zif[kMaxLayers];  // interface depths (interior-relative cells), ascending
  double vlayer[kMaxLayers];
  double bc[kMaxBlobs][3], bs2[kMaxBlobs], ba[kMaxBlobs];  // centre, 2 sigma^2, amplitude
  double pert, dt_over_h, anomaly, ac[3], as2;
};

uint64_t splitmix64_host(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01_host(uint64_t &state) { return (double)(splitmix64_host(state) >> 11) * 0x1.0p-53; }

__device__ __forceinline__ double model_speed(const ModelParams &m, double x, double y, double z) {
  int l = 0;
  for (int q = 0; q < m.nlayers - 1; ++q) l += z >= m.zif[q] ? 1 : 0;
  double v = m.vlayer[l];
  double g = 0.0;
  for (int b = 0; b < m.nblobs; ++b) {
    const double dx = x - m.bc[b][0], dy = y - m.bc[b][1], dz = z - m.bc[b][2];
    g += m.ba[b] * exp(-(dx * dx + dy * dy + dz * dz) / m.bs2[b]);
  }
  g = fmin(1.0, fmax(-1.0, g));
  v *= 1.0 + m.pert * g;
  if (m.anomaly != 0.0) {
    const double dx = x - m.ac[0], dy = y - m.ac[1], dz = z - m.ac[2];
    v *= 1.0 + m.anomaly * exp(-(dx * dx + dy * dy + dz * dz) / m.as2);
  }
  return v;
}

template <typename C>
__device__ __forceinline__ void store_roc2(C *p, double r) {
  *p = (C)r;
}
template <>
__device__ __forceinline__ void store_roc2<__hip_bfloat16>(__hip_bfloat16 *p, double r) {
  *p = __float2bfloat16((float)r);
}

// roc2 over the whole halo-inclusive field; halo cells take the speed of the
// nearest interior cell (kernel 0 reads roc2 at interior cells only).
template <typename C>
__global__ void model_kernel(C *roc2, const ModelParams m, long long total) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int nnx = m.nx + 2 * m.R;
  const long long row = t / nnx;
  const int i = (int)(t - row * nnx);
  const int j = (int)(row % m.rows);
  const int k = (int)(row / m.rows);
  const double x = (double)min(max(i - m.R, 0), m.nx - 1);
  const double y = (double)min(max(j - m.R, 0), m.ny - 1);
  const double z = (double)min(max(k - m.R, 0), m.nz - 1);
  const double c = model_speed(m, x, y, z) * m.dt_over_h;
  double r = c * c;
  if (m.round_bf16) r = (double)__bfloat162float(__float2bfloat16((float)r));
  store_roc2<C>(roc2 + row * m.pitch + i, r);
}

// bf16 roc2 -> the field type, same layout
template <typename T>
__global__ void roc2_convert_bf16_kernel(const __hip_bfloat16 *src, T *dst, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (T)__bfloat162float(src[t]);
}

}  // namespace

using gr::Problem;

extern "C" int gr_init_roc2_model(gr_handle *h, const gr_vmodel *vm) {
  if (!h || !vm) return gr::set_error(GR_EINVAL, "gr_init_roc2_model: null argument");
  const Problem &P = gr_internal_problem(h);
  if (P.kernel != 0 || !P.roc2)
    return gr::set_error(GR_EINVAL, "gr_init_roc2_model: the problem has no roc2 field (kernel 0 without "
                                    "GR_ROC2_SCALAR)");
  if (vm->nlayers < 1 || vm->nlayers > kMaxLayers || vm->nblobs < 0 || vm->nblobs > kMaxBlobs)
    return gr::set_errorf(GR_EINVAL, "gr_init_roc2_model: nlayers must be 1..%d and nblobs 0..%d", kMaxLayers,
                          kMaxBlobs);
  if (!(vm->vmin > 0) || !(vm->vmax >= vm->vmin) || !(vm->dt_over_h > 0) || vm->pert < 0 || vm->pert >= 1 ||
      vm->anomaly <= -1 || (vm->anomaly != 0 && !(vm->anomaly_sigma > 0)))
    return gr::set_error(GR_EINVAL, "gr_init_roc2_model: bad speeds, perturbation or time step");
  ModelParams m{};
  m.nlayers = vm->nlayers;
  m.nblobs = vm->nblobs;
  m.round_bf16 = vm->round_bf16 ? 1 : 0;
  m.nx = P.nx;
  m.ny = P.ny;
  m.nz = P.nz;
  m.R = P.R;
  m.pitch = P.pitch;
  m.rows = P.rows;
  uint64_t st = vm->seed * 0xD1B54A32D192ED03ull + 0x5851F42D4C957F2Dull;
  std::vector<double> zif;
  for (int q = 0; q < m.nlayers - 1; ++q) zif.push_back((0.05 + 0.9 * u01_host(st)) * P.nz);
  std::sort(zif.begin(), zif.end());
  for (int q = 0; q < m.nlayers - 1; ++q) m.zif[q] = zif[q];
  for (int l = 0; l < m.nlayers; ++l)
    m.vlayer[l] = m.nlayers == 1 ? vm->vmin : vm->vmin + (vm->vmax - vm->vmin) * l / (m.nlayers - 1);
  const double nmin = (double)std::min(P.nx, std::min(P.ny, P.nz));
  for (int b = 0; b < m.nblobs; ++b) {
    m.bc[b][0] = u01_host(st) * P.nx;
    m.bc[b][1] = u01_host(st) * P.ny;
    m.bc[b][2] = u01_host(st) * P.nz;
    const double sig = nmin * (1.0 / 12 + (1.0 / 5 - 1.0 / 12) * u01_host(st));
    m.bs2[b] = 2 * sig * sig;
    m.ba[b] = 2 * u01_host(st) - 1;
  }
  m.pert = vm->pert;
  m.dt_over_h = vm->dt_over_h;
  m.anomaly = vm->anomaly;
  for (int d = 0; d < 3; ++d) m.ac[d] = vm->anomaly_center[d];
  m.as2 = 2 * vm->anomaly_sigma * vm->anomaly_sigma;
  const long long total = (long long)(P.nx + 2 * P.R) * P.rows * P.planes;
  const int bs = 256;
  const unsigned nb = (unsigned)((total + bs - 1) / bs);
  hipStream_t s = gr_internal_stream(h);
  if (P.flags & GR_COEF_BF16)
    hipLaunchKernelGGL(model_kernel<__hip_bfloat16>, dim3(nb), dim3(bs), 0, s, (__hip_bfloat16 *)P.roc2, m, total);
  else if (P.dtype == GR_F64)
    hipLaunchKernelGGL(model_kernel<double>, dim3(nb), dim3(bs), 0, s, (double *)P.roc2, m, total);
  else
    hipLaunchKernelGGL(model_kernel<float>, dim3(nb), dim3(bs), 0, s, (float *)P.roc2, m, total);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return gr::set_errorf(GR_EHIP, "gr_init_roc2_model: %s", hipGetErrorString(e));
  return GR_OK;
}

extern "C" int gr_download_roc2(gr_handle *h, const int ld[3], void *U3) {
  if (!h || !ld || !U3) return gr::set_error(GR_EINVAL, "gr_download_roc2: null argument");
  const Problem &P = gr_internal_problem(h);
  if (!P.roc2) return gr::set_error(GR_EINVAL, "gr_download_roc2: the problem has no roc2 field");
  if (ld[0] < P.nx + 2 * P.R || ld[1] != P.ny + 2 * P.R || ld[2] != P.nz + 2 * P.R)
    return gr::set_error(GR_EINVAL, "gr_download_roc2: ldomain_shape does not fit the grid");
  const size_t es = gr::elem_size(P.dtype);
  const size_t width = (size_t)(P.nx + 2 * P.R);
  hipStream_t s = gr_internal_stream(h);
  const void *src = P.roc2;
  void *tmp = nullptr;
  hipError_t e = hipSuccess;
  if (P.flags & GR_COEF_BF16) {  // widen to the field type first
    const long long n = P.field_elems;
    e = hipMalloc(&tmp, (size_t)n * es);
    if (e == hipSuccess) {
      const __hip_bfloat16 *b = (const __hip_bfloat16 *)P.roc2 - P.xoff;
      const unsigned nb = (unsigned)((n + 255) / 256);
      if (P.dtype == GR_F64)
        hipLaunchKernelGGL(roc2_convert_bf16_kernel<double>, dim3(nb), dim3(256), 0, s, b, (double *)tmp, n);
      else
        hipLaunchKernelGGL(roc2_convert_bf16_kernel<float>, dim3(nb), dim3(256), 0, s, b, (float *)tmp, n);
      e = hipGetLastError();
      src = (const char *)tmp + (size_t)P.xoff * es;
    }
  }
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(U3, (size_t)ld[0] * es, src, (size_t)P.pitch * es, width * es, (size_t)P.rows * P.planes,
                         hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess) return gr::set_errorf(GR_EHIP, "gr_download_roc2: %s", hipGetErrorString(e));
  return GR_OK;
}
