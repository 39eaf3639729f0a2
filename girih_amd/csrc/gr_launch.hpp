// gr_launch.hpp -- host-side dispatch from a problem's kernel id, arithmetic mode
// and launch configuration to one instantiation of sweep_kernel.  The
// instantiations are split by element type over three translation units
// (gr_launch_f64.hip, gr_launch_f32.hip, gr_launch_bf16.hip) so they compile in
// parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <type_traits>

#include "gr_internal.hpp"
#include "gr_sweep.hpp"

namespace gr {

hipError_t launch_sweep_f64(int kid, bool fast, bool rocs, const SweepArgs<double, double> &a, const LaunchCfg &lc,
                            hipStream_t st, bool nt);
hipError_t launch_sweep_f32(int kid, bool fast, bool rocs, const SweepArgs<float, float> &a, const LaunchCfg &lc,
                            hipStream_t st, bool nt);
hipError_t launch_sweep_bf16(int kid, bool fast, bool rocs, const SweepArgs<float, __hip_bfloat16> &a,
                             const LaunchCfg &lc, hipStream_t st, bool nt);

#ifdef GR_LAUNCH_IMPL
template <typename T, typename C, int KID, bool FAST, bool ROCS, bool NT, int BY, int RY, int MINW, bool RING = false,
          int BX = 1, bool PF = false>
inline hipError_t launch_cfg(const SweepArgs<T, C> &a0, const LaunchCfg &lc, hipStream_t st) {
  SweepArgs<T, C> a = a0;
  const int TX = 64 * BX, TY = BY * RY;
  const int bx = a.g.xe - a.g.xb, by = a.g.ye - a.g.yb, bz = a.g.ze - a.g.zb;
  a.tiles_x = (bx + TX - 1) / TX;
  a.tiles_y = (by + TY - 1) / TY;
  a.zchunk = lc.zchunk > 0 ? std::min(lc.zchunk, bz) : bz;
  a.nzc = (bz + a.zchunk - 1) / a.zchunk;
  a.xcd_order = lc.xcd_order;
  const int nb = a.tiles_x * a.tiles_y * a.nzc;
  // Buffer-resource plane accesses where they measured faster: FP32 kernel 0 with a
  // roc2 field (bf16 roc2 at 1024^3: 307 -> 353 GLUP/s).  Scalar-roc2 FP32 (-2%) and
  // FP64 (-11%) sweeps keep SGPR-base global loads.
  constexpr bool BUF = KID == 0 && !ROCS && std::is_same<T, float>::value;
  hipLaunchKernelGGL((sweep_kernel<T, C, KID, BY, RY, FAST, ROCS, MINW, RING, kFastND, BX, NT, BUF, PF && KID == 0>), dim3(nb),
                     dim3(TX, BY), 0, st, a);
  return hipGetLastError();
}

// The LDS plane ring holds R+2 planes of the tile plus halo; where that exceeds the
// 160 KB of LDS of a CU the variant falls back to the double-buffered tile.
template <typename T, int KID, int BX, int BY, int RY>
constexpr bool ring_fits() {
  constexpr int R = kradius(KID);
  return (size_t)(R + 2) * (BY * RY + 2 * R) * (64 * BX + 2 * R) * sizeof(T) <= 160 * 1024;
}

template <typename T, typename C, int KID, bool FAST, bool ROCS, bool NT>
inline hipError_t launch_variant(const SweepArgs<T, C> &a, const LaunchCfg &lc, hipStream_t st) {
  constexpr int BY = Tile<T, KID>::BY, RY = Tile<T, KID>::RY;
  constexpr int W = std::is_same<T, double>::value ? 3 : 4;
  if constexpr (KID == 0 || KID == 1 || KID == 5) {
    switch (lc.variant) {
      case 1: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 4, 0>(a, lc, st);
      case 2: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 2, 0>(a, lc, st);
      case 3: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 4, W>(a, lc, st);
      case 4: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, W>(a, lc, st);
      case 5: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 4, W, true>(a, lc, st);
      case 6: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 2, 0, true>(a, lc, st);
      case 7: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, W, true>(a, lc, st);
      case 8: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, 0, false, 2>(a, lc, st);
      case 9: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 4, 0, false, 2>(a, lc, st);
      case 10: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, 0, true, 2>(a, lc, st);
      case 11: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, W, false, 2>(a, lc, st);
      case 12: return launch_cfg<T, C, KID, FAST, ROCS, NT, 2, 4, 0, true, 2>(a, lc, st);
      case 13: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 2, 0, true, 2>(a, lc, st);
      case 14: return launch_cfg<T, C, KID, FAST, ROCS, NT, 4, 4, W, true, 2>(a, lc, st);
      case 15: return launch_cfg<T, C, KID, FAST, ROCS, NT, 2, 4, 0, ring_fits<T, KID, 4, 2, 4>(), 4>(a, lc, st);
      case 16: return launch_cfg<T, C, KID, FAST, ROCS, NT, 8, 2, 0, true, 2, true>(a, lc, st);
      default: break;
    }
  }
  return launch_cfg<T, C, KID, FAST, ROCS, NT, BY, RY, 0>(a, lc, st);
}

// Non-temporal hints (NT) where the problem's working set exceeds the caches.
template <typename T, typename C, int KID, bool FAST, bool ROCS>
inline hipError_t launch_sweep_t(const SweepArgs<T, C> &a, const LaunchCfg &lc, hipStream_t st, bool nt) {
  return nt ? launch_variant<T, C, KID, FAST, ROCS, true>(a, lc, st) : launch_variant<T, C, KID, FAST, ROCS, false>(a, lc, st);
}

template <typename T, typename C>
inline hipError_t launch_sweep_k(int kid, bool fast, bool rocs, const SweepArgs<T, C> &a,
                                 const LaunchCfg &lc, hipStream_t st, bool nt = false) {
  switch (kid) {
    case 0:
      if constexpr (std::is_same<T, float>::value) {
        if (fast) {
          if (rocs) return launch_sweep_t<T, C, 0, true, true>(a, lc, st, nt);
          return launch_sweep_t<T, C, 0, true, false>(a, lc, st, nt);
        }
      }
      if (rocs) return launch_sweep_t<T, C, 0, false, true>(a, lc, st, nt);
      return launch_sweep_t<T, C, 0, false, false>(a, lc, st, nt);
    case 1: return launch_sweep_t<T, C, 1, false, false>(a, lc, st, nt);
    case 2: return launch_sweep_t<T, C, 2, false, false>(a, lc, st, nt);
    case 3: return launch_sweep_t<T, C, 3, false, false>(a, lc, st, nt);
    case 4: return launch_sweep_t<T, C, 4, false, false>(a, lc, st, nt);
    case 5: return launch_sweep_t<T, C, 5, false, false>(a, lc, st, nt);
  }
  return hipErrorInvalidValue;
}

#endif  // GR_LAUNCH_IMPL

}  // namespace gr
