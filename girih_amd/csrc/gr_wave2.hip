// gr_wave2.hip -- the single-grid kernel of the two-step kernel-0 pass (the tile
// body, Wave2Tile, is in gr_wave2.hpp and shared with the batched kernel), its
// launch configuration and the host entry wave2_pass.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <type_traits>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"
#include "gr_wave2.hpp"

namespace gr {

template <typename T, typename C, bool FAST, bool ROCS, bool STRM, int BX, int BY, int RY2>
__global__ __launch_bounds__(64 * BX * BY) void wave2_kernel(const W2Args<T, C> a) {
  using WT = Wave2Tile<T, C, FAST, ROCS, STRM, BX, BY, RY2>;
  __shared__ typename WT::Lds lds;
  const int ntile = a.tiles_x * a.tiles_y;
  const int lb = tile_order(blockIdx.x, ntile * a.nzc, ntile, a.xcd_order);
  WT::run(a, lb % ntile, lb / ntile, lds);
}

template <typename T, typename C, bool FAST, bool ROCS, bool STRM, int BX, int BY, int RY2>
static hipError_t launch_w2(W2Args<T, C> a, hipStream_t st) {
  constexpr int TX = 64 * BX, TY = BY * RY2;
  if (g_fused_tile_probe) {
    g_fused_tile_probe[0] = TX;
    g_fused_tile_probe[1] = TY;
    return hipSuccess;
  }
  a.tiles_x = (a.nx + TX - 1) / TX;
  a.tiles_y = (a.ny + TY - 1) / TY;
  a.zchunk = (a.zchunk <= 0 || a.zchunk > a.nz) ? a.nz : a.zchunk;
  a.nzc = (a.nz + a.zchunk - 1) / a.zchunk;
  hipLaunchKernelGGL((wave2_kernel<T, C, FAST, ROCS, STRM, BX, BY, RY2>), dim3(a.tiles_x * a.tiles_y * a.nzc),
                     dim3(64 * BX * BY), 0, st, a);
  return hipGetLastError();
}

// Output tiles of 64 x 16 cells and 512 lanes (FP32: 64 KB LDS, two workgroups per
// CU; FP64: 128 KB).  In FP32 the 128 x 16 tile of 1024 lanes (one workgroup per CU,
// the 192-word LDS pitch) measured 2-4% slower; a 128 x 8 FP64 tile of 1024 lanes
// spilled its queues and measured 25% slower.
template <typename T, typename C, bool FAST>
static hipError_t w2_dispatch(bool rocs, bool strm, const W2Args<T, C> &a, hipStream_t st) {
  constexpr int BX = 1;
  if constexpr (std::is_same<T, C>::value) {  // a scalar roc2 has no storage type of its own
    if (rocs) {
      if (strm) return launch_w2<T, C, FAST, true, true, BX, 8, 2>(a, st);
      return launch_w2<T, C, FAST, true, false, BX, 8, 2>(a, st);
    }
  }
  if (strm) return launch_w2<T, C, FAST, false, true, BX, 8, 2>(a, st);
  return launch_w2<T, C, FAST, false, false, BX, 8, 2>(a, st);
}

template <typename T, typename C>
W2Args<T, C> w2_args(const Problem &P, const void *X, const void *Y, void *ZO, void *ZN, const void *src,
                     long long src_idx, const LaunchCfg &lc) {
  const long long sz = (long long)P.pitch * P.rows;
  W2Args<T, C> a{};
  a.P = (const T *)X;
  a.V = (const T *)Y;
  a.W = (T *)ZO;
  a.Y = (T *)ZN;
  a.roc2 = (const C *)P.roc2;
  for (int r = 0; r < 5; ++r) a.c[r] = (T)P.c[r];
  a.csum = (T)fast_csum(P.c);
  a.roc2s = (T)P.roc2s;
  a.sz = sz;
  a.sy = P.pitch;
  a.nx = P.nx;
  a.ny = P.ny;
  a.nz = P.nz;
  a.xcd_order = lc.xcd_order;
  a.zchunk = lc.zchunk;
  a.src = (const T *)src;
  a.src_k = src_idx < 0 ? -1 : (int)(src_idx / sz);
  a.src_off = src_idx < 0 ? 0 : (int)(src_idx % sz);
  return a;
}

// One pass of two kernel-0 steps: (X, Y) = (u^(n-1), u^n) -> (ZO, ZN) = (u^(n+1), u^(n+2)).
hipError_t wave2_pass(const Problem &P, const void *X, const void *Y, void *ZO, void *ZN, const void *src,
                      long long src_idx, const LaunchCfg &lc, hipStream_t st) {
  if (P.kernel != 0) return hipErrorInvalidValue;
  const bool fast = (P.flags & GR_FAST) != 0, rocs = (P.flags & GR_ROC2_SCALAR) != 0;
  const bool strm = streaming_sweep(P) != 0;
  if (P.dtype == GR_F64) {
    if (P.flags & GR_COEF_BF16) return hipErrorInvalidValue;
    return w2_dispatch<double, double, false>(rocs, strm, w2_args<double, double>(P, X, Y, ZO, ZN, src, src_idx, lc),
                                              st);
  }
  if ((P.flags & GR_COEF_BF16) && !rocs) {  // a bf16 roc2 field (a scalar roc2 has no storage type)
    const W2Args<float, __hip_bfloat16> a = w2_args<float, __hip_bfloat16>(P, X, Y, ZO, ZN, src, src_idx, lc);
    if (fast) return w2_dispatch<float, __hip_bfloat16, true>(false, strm, a, st);
    return w2_dispatch<float, __hip_bfloat16, false>(false, strm, a, st);
  }
  const W2Args<float, float> a = w2_args<float, float>(P, X, Y, ZO, ZN, src, src_idx, lc);
  if (fast) return w2_dispatch<float, float, true>(rocs, strm, a, st);
  return w2_dispatch<float, float, false>(rocs, strm, a, st);
}

}  // namespace gr
