// gr_tblock.hip -- temporally blocked sweeps: TL time steps per HBM pass.
//
// The GPU analogue of girih's diamond + wavefront temporal blocking
// (src/kernels/diamond_ts.c, stencils_1wf.ic:19-82): instead of diamonds in y-t
// and a wavefront in z run by a CPU thread, a workgroup owns an xy tile and
// streams z once, computing TL time levels in a z pipeline.  Level t trails level
// t-1 by R planes -- the `kt -= NHALO` of stencils_1wf.ic:77 -- so every level
// finds the planes it needs already computed on chip:
//
//   step k0:  level 0 = plane k0 of the input enters the pipeline,
//             level t computes plane k0 - R*t   (t = 1 .. TL)
//
// Level t is computed over the tile plus an xy margin of R*(TL - t) cells (the
// overlapped ghost zone that makes the tile independent of its neighbours).  A
// lane owns columns of the level-1 region; for each level it keeps a register
// queue of the 2R+1 most recent planes (the z neighbours of the next level), and
// the current plane of each level is staged in LDS for the xy neighbours of the
// next level.
// Only the input field(s) are read from HBM and only the tile's output planes are
// written, so a pass moves (in + out) words per cell for TL updates.
//
// Boundary cells (girih's halo, never updated) keep their input value at every
// level; the host enables this path only when U1 and U2 hold the same halo.
//
// This kernel serves kernel 5 (seven coefficient arrays, FP64, FP32 or bf16-stored):
// level t = f(level t-1), input Y, output ZN.  Kernel 1 has a leaner pass in gr_fused1.hip and kernel 0
// (second order in time) a two-step pass in gr_wave2.hip.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"
#include "gr_sweep.hpp"

namespace gr {

// C: storage type of the coefficient arrays (T, or bf16 widened exactly to FP32)
template <typename T, typename C>
struct TbArgs {
  const T *Y;   // level 0
  T *ZN;        // level TL
  const C *coef;
  long long cstride;
  long long sy, sz;
  int nx, ny, nz;  // interior; boxes are [R, n+R)
  int tiles_x, tiles_y, xcd_order;
  int zchunk, nzc;  // output planes per workgroup, chunks per tile column
  const T *src;    // source values of this pass (level t adds src[t-1]); src_k < 0: none
  int src_k, src_off;
};

// CP: level t+1 takes plane p's coefficients from level t one step earlier (same lane
// and column) instead of re-reading them; costs TL x 7 registers per column, or TL x 4
// with bf16 coefficients, which are handed on as stored, two to a register.
template <typename T, typename C, int KID, int NT, int TXO, int TYO, int TL, bool CP>
__global__ __launch_bounds__(NT) void tb_kernel(const TbArgs<T, C> a) {
  static_assert(KID == 5, "kernel 5 only");
  constexpr int R = kradius(KID);
  constexpr int Q = 2 * R + 1;
  constexpr int W0 = TXO + 2 * R * TL, H0 = TYO + 2 * R * TL;          // level-0 region
  constexpr int W1 = TXO + 2 * R * (TL - 1), H1 = TYO + 2 * R * (TL - 1);  // level-1 region
  constexpr int NCOL = (W1 * H1 + NT - 1) / NT;                            // columns per lane
  constexpr int NH = W0 * H0 - W1 * H1;                                    // level-0 halo ring
  constexpr int HS = (NH + NT - 1) / NT;
  constexpr int PL = W0 * H0;  // one LDS plane (region-0 coordinates for every level)
  __shared__ T lds[TL][PL];

  const int ntile = a.tiles_x * a.tiles_y;
  const int lb = tile_order(blockIdx.x, ntile * a.nzc, ntile, a.xcd_order);
  const int tile = lb % ntile;
  const int tx0 = R + (tile % a.tiles_x) * TXO;  // tile origin (girih interior indices)
  const int ty0 = R + (tile / a.tiles_x) * TYO;
  // output planes [zs, ze) of this workgroup; level t is needed on
  // [zs - R*(TL-t), ze + R*(TL-t)), the z analogue of the xy ghost margin
  const int zs = R + (lb / ntile) * a.zchunk;
  const int ze = min(zs + a.zchunk, a.nz + R);
  const int nnx = a.nx + 2 * R, nny = a.ny + 2 * R, nnz = a.nz + 2 * R;
  const int tid = threadIdx.x;

  // lane columns in region-1 coordinates
  unsigned off[NCOL];   // in-plane byte offset (clamped into the field)
  int lofs[NCOL];       // LDS index (region-0 coordinates)
  int cxv[NCOL], cyv[NCOL];
  bool inner[NCOL], outc[NCOL], colv[NCOL];
#pragma unroll
  for (int j = 0; j < NCOL; ++j) {
    const int c = min(tid + j * NT, W1 * H1 - 1);
    const int cx = c % W1, cy = c / W1;
    colv[j] = tid + j * NT < W1 * H1;  // the last lane slots may be spare
    cxv[j] = colv[j] ? cx : -1000000;
    cyv[j] = cy;
    const int gx = tx0 - R * (TL - 1) + cx, gy = ty0 - R * (TL - 1) + cy;
    const int gxc = min(max(gx, 0), nnx - 1), gyc = min(max(gy, 0), nny - 1);
    off[j] = (unsigned)(gyc * (int)a.sy + gxc) * (unsigned)sizeof(T);
    lofs[j] = (cy + R) * W0 + (cx + R);
    inner[j] = gx >= R && gx < a.nx + R && gy >= R && gy < a.ny + R;
    // output: tile cells inside the interior
    outc[j] = (tid + j * NT < W1 * H1) && cx >= R * (TL - 1) && cx < R * (TL - 1) + TXO &&
              cy >= R * (TL - 1) && cy < R * (TL - 1) + TYO && inner[j];
  }
  // level-0 halo ring slots
  unsigned hoff[HS];
  int hl[HS];
  bool hw[HS];
#pragma unroll
  for (int s = 0; s < HS; ++s) {
    const int e = min(tid + s * NT, NH - 1);
    hw[s] = tid + s * NT < NH;
    // ring: R*TL... enumerate region 0 cells that are not in region 1
    int lx, ly;
    {
      const int top = R * W0;  // R rows above region 1
      if (e < top) {
        ly = e / W0;
        lx = e % W0;
      } else if (e < 2 * top) {
        ly = R + H1 + (e - top) / W0;
        lx = (e - top) % W0;
      } else {
        const int d = e - 2 * top, side = d / (R * H1), f = d - side * R * H1;
        ly = R + f / R;
        lx = (f % R) + side * (R + W1);
      }
    }
    const int gx = tx0 - R * TL + lx, gy = ty0 - R * TL + ly;
    const int gxc = min(max(gx, 0), nnx - 1), gyc = min(max(gy, 0), nny - 1);
    hoff[s] = (unsigned)(gyc * (int)a.sy + gxc) * (unsigned)sizeof(T);
    hl[s] = ly * W0 + lx;
  }

  auto planeY = [&](int k) { return a.Y + (long long)k * a.sz; };

  // queues: qv[t][j][slot], level t plane p in slot (p mod Q)
  T qv[TL][NCOL][Q];
#pragma unroll
  for (int t = 0; t < TL; ++t)
#pragma unroll
    for (int j = 0; j < NCOL; ++j)
#pragma unroll
      for (int i = 0; i < Q; ++i) qv[t][j][i] = T(0);

  const int kfirst = zs - R * TL, klast = ze + R * TL;  // k0 range of this chunk
  // prefetch level-0 plane kfirst (only planes inside [0, nnz) exist)
  T n0[NCOL], nh[HS];
  // interior columns of plane k (queue entry) and the halo ring of plane k - R
  // (the level-0 plane that level 1 centres on when plane k enters the pipeline)
  auto fetch0 = [&](int k) {
    const bool ok = k >= 0 && k < nnz, okr = k - R >= 0 && k - R < nnz;
    const T *vp = planeY(min(max(k, 0), nnz - 1));
    const T *rp = planeY(min(max(k - R, 0), nnz - 1));
#pragma unroll
    for (int j = 0; j < NCOL; ++j) n0[j] = ok ? ldb<T>(vp, off[j]) : T(0);
#pragma unroll
    for (int s = 0; s < HS; ++s) nh[s] = okr ? ldb<T>(rp, hoff[s]) : T(0);
  };
  fetch0(kfirst);

  // cs[t-1][j]: the coefficients level t uses in this step.  Level t+1 is live at
  // (step k0+1, column j) only where level t was live at (k0, j) on the same plane p =
  // k0 - t: level regions shrink with t and level t+1 starts R planes later.  bf16
  // coefficients are handed on as stored, two to a register (PK).
  constexpr bool PK = CP && std::is_same<C, __hip_bfloat16>::value;
  constexpr int CW = PK ? 4 : 7;  // registers per column and level
  using CS = typename std::conditional<PK, unsigned, T>::type;
  CS cs[CP ? TL : 1][NCOL][CW];

  for (int kb = kfirst; kb < klast; kb += Q) {
#pragma unroll
    for (int ph = 0; ph < Q; ++ph) {
      const int k0 = kb + ph;
      if (k0 >= klast) break;
      if constexpr (CP) {
#pragma unroll
        for (int t = TL - 1; t >= 1; --t)
#pragma unroll
          for (int j = 0; j < NCOL; ++j)
#pragma unroll
            for (int m = 0; m < CW; ++m) cs[t][j][m] = cs[t - 1][j][m];
      }
      // slot of plane p: ((p mod Q) + Q) mod Q; k0 - kfirst = kb - kfirst + ph and
      // (kb - kfirst) is a multiple of Q, so plane k0 - R*t sits in slot
      // (ph - R*t) mod Q relative to kfirst's slot 0.
      // ---- level 0: plane k0 enters the queue; LDS gets plane k0 - R, the centre
      // plane of level 1 in this step (from the queue, R steps old) and its halo ring
#pragma unroll
      for (int j = 0; j < NCOL; ++j) qv[0][j][ph] = n0[j];
      {
        T *L0 = lds[0];
#pragma unroll
        for (int j = 0; j < NCOL; ++j)
          if (colv[j]) L0[lofs[j]] = qv[0][j][((ph - R) % Q + Q) % Q];
#pragma unroll
        for (int s = 0; s < HS; ++s)
          if (hw[s]) L0[hl[s]] = nh[s];
      }
      fetch0(k0 + 1);
      __syncthreads();

#pragma unroll
      for (int t = 1; t <= TL; ++t) {
        const int p = k0 - R * t;  // plane computed by level t
        const bool live = p >= 0 && p < nnz && p >= zs - R * (TL - t);  // earlier planes feed nothing
        const bool zin = p >= R && p < a.nz + R;
        const int sc = ((ph - R * t) % Q + Q) % Q;  // slot of plane p in level t-1's queue
        T val[NCOL];
        if (live) {
          const T *Lp = lds[t - 1];
          const T *yp = planeY(p);
#pragma unroll
          for (int j = 0; j < NCOL; ++j) {
            // region of level t in region-1 coordinates: [R(t-1), W1-R(t-1))
            const bool inreg = cxv[j] >= R * (t - 1) && cxv[j] < W1 - R * (t - 1) && cyv[j] >= R * (t - 1) &&
                               cyv[j] < H1 - R * (t - 1);
            if (!inreg) {
              val[j] = T(0);
              continue;
            }
            T out;
            if (inner[j] && zin) {
              const T *q = qv[t - 1][j];
              auto zn = [&](int d) { return q[((sc + d) % Q + Q) % Q]; };
              auto xn = [&](int d) { return Lp[lofs[j] + d]; };
              auto yn = [&](int d) { return Lp[lofs[j] + d * W0]; };
              T zq[Q];
#pragma unroll
              for (int d = -R; d <= R; ++d) zq[R + d] = zn(d);
              const T v0 = zq[R];
              {  // stencils.c:193-201; coefficients re-read by every level stay cached
                const C *cp = a.coef + (long long)p * a.sz + off[j] / sizeof(T);
                if constexpr (PK) {  // the bits of a bf16 value are the top half of its FP32 value
                  if (t == 1) {
                    const unsigned short *cb = reinterpret_cast<const unsigned short *>(cp);
                    unsigned short h[7];
#pragma unroll
                    for (int m = 0; m < 7; ++m) h[m] = cb[m * a.cstride];
#pragma unroll
                    for (int w = 0; w < 3; ++w) cs[0][j][w] = (unsigned)h[2 * w] | ((unsigned)h[2 * w + 1] << 16);
                    cs[0][j][3] = h[6];
                  }
                  const unsigned *w = cs[t - 1][j];
                  T c[7];
#pragma unroll
                  for (int m = 0; m < 7; ++m)
                    c[m] = __uint_as_float((m & 1) ? (w[m >> 1] & 0xffff0000u) : (w[m >> 1] << 16));
                  out = c[0] * v0 + c[1] * xn(-1) + c[2] * xn(1) + c[3] * yn(-1) + c[4] * yn(1) + c[5] * zq[0] +
                        c[6] * zq[2];
                } else if constexpr (CP) {
                  if (t == 1) {
#pragma unroll
                    for (int m = 0; m < 7; ++m) cs[0][j][m] = ld_coef<T, C, false>(cp + m * a.cstride);
                  }
                  const T *c = cs[CP ? t - 1 : 0][j];
                  out = c[0] * v0 + c[1] * xn(-1) + c[2] * xn(1) + c[3] * yn(-1) + c[4] * yn(1) + c[5] * zq[0] +
                        c[6] * zq[2];
                } else {
                  auto cf = [&](int m) { return ld_coef<T, C, false>(cp + m * a.cstride); };
                  out = cf(0) * v0 + cf(1) * xn(-1) + cf(2) * xn(1) + cf(3) * yn(-1) + cf(4) * yn(1) +
                        cf(5) * zq[0] + cf(6) * zq[2];
                }
              }
              if (a.src_k >= 0 && p == a.src_k && off[j] == (unsigned)a.src_off * (unsigned)sizeof(T)) out = out + a.src[t - 1];
            } else {
              out = ldb<T>(yp, off[j]);  // girih's halo: never updated
            }
            val[j] = out;
          }
        } else {
#pragma unroll
          for (int j = 0; j < NCOL; ++j) val[j] = T(0);
        }
        // level t: plane p into its queue; the LDS gets its plane p - R (the centre
        // plane of level t+1 in this step), taken from the queue
        if (t < TL) {
          const int sw = ((ph - R * t) % Q + Q) % Q;
          const int sl = ((ph - R * t - R) % Q + Q) % Q;
          T *Lt = lds[t < TL ? t : 0];
#pragma unroll
          for (int j = 0; j < NCOL; ++j) qv[t < TL ? t : 0][j][sw] = val[j];
#pragma unroll
          for (int j = 0; j < NCOL; ++j)
            if (colv[j]) Lt[lofs[j]] = qv[t < TL ? t : 0][j][sl];
        }
        // output: level TL into ZN -- planes of the box
        if (t == TL && live && zin && p >= zs && p < ze) {
#pragma unroll
          for (int j = 0; j < NCOL; ++j)
            if (outc[j]) stb<T>(a.ZN + (long long)p * a.sz, off[j], val[j]);
        }
        if (t < TL) __syncthreads();
      }
    }
  }
}

template <typename T, typename C, int KID, int NT, int TXO, int TYO, int TL, bool CP = false>
static hipError_t launch_tb(const TbArgs<T, C> &a0, hipStream_t st) {
  if (g_fused_tile_probe) {
    g_fused_tile_probe[0] = TXO;
    g_fused_tile_probe[1] = TYO;
    return hipSuccess;
  }
  TbArgs<T, C> a = a0;
  a.tiles_x = (a.nx + TXO - 1) / TXO;
  a.tiles_y = (a.ny + TYO - 1) / TYO;
  a.zchunk = (a.zchunk <= 0 || a.zchunk > a.nz) ? a.nz : a.zchunk;
  a.nzc = (a.nz + a.zchunk - 1) / a.zchunk;
  hipLaunchKernelGGL((tb_kernel<T, C, KID, NT, TXO, TYO, TL, CP>), dim3(a.tiles_x * a.tiles_y * a.nzc), dim3(NT), 0, st, a);
  return hipGetLastError();
}

// Fused depths available per kernel (0 = not available)
int tb_depths(int kid, int dtype, int *out, int max) {
  int n = 0;
  auto add = [&](int d) {
    if (n < max) out[n++] = d;
  };
  (void)dtype;
  if (kid == 1) {  // gr_fused1.hip
    for (int d = 2; d <= 6; ++d) add(d);
  } else if (kid == 5) {  // tb_kernel
    add(2);
    add(4);
    add(6);
  } else if (kid == 0) {
    add(2);  // gr_wave2.hip
  }
  return n;
}

// Tiles per kernel and precision, sized so the per-lane queues (levels x columns
// x (2R+1) words) and the coefficient hand-off stay in registers without spilling.
template <typename T, typename C>
static hipError_t tb_dispatch(int kid, int tl, const TbArgs<T, C> &a, hipStream_t st) {
  constexpr bool F32 = std::is_same<T, float>::value;
  constexpr bool B16 = std::is_same<C, __hip_bfloat16>::value;
  if (kid == 5) {
    // depth 2, coefficient hand-off. FP32: 896 lanes give 3 columns per lane (66 x 34
    // level-1 cells), 86 VGPRs and 5 waves per SIMD, against 5 columns, 139 VGPRs and
    // 3 waves with 512 lanes.  FP64: 640 lanes give 2 columns per lane (66 x 18
    // level-1 cells), 112 VGPRs and 4 waves per SIMD, against 3 columns, 156 VGPRs and
    // 3 waves with 512 lanes
    if (tl == 2) {
      if constexpr (F32) return launch_tb<T, C, 5, 896, 64, 32, 2, true>(a, st);
      else return launch_tb<T, C, 5, 640, 64, 16, 2, true>(a, st);
    }
    // depth 4, coefficient hand-off on 1024 lanes (512^3, 256-plane chunks, one box):
    //   bf16 coefficients, 64 x 32 tile (3 columns per lane, 117 VGPRs): 388 GLUP/s,
    //   against 274 re-reading them on 512 lanes;
    //   FP32 coefficients, 64 x 16 tile (2 columns, 110 VGPRs): 284 against 206 (896
    //   lanes re-reading them) and 221 at depth 2; a 64 x 32 tile spills;
    //   FP64, 64 x 8 tile (1 column, 110 VGPRs): 128 against 79 re-reading them on 512
    //   lanes and 104 at depth 2; 896 lanes (2 columns) spill.
    if (tl == 4) {
      if constexpr (B16) return launch_tb<T, C, 5, 1024, 64, 32, 4, true>(a, st);
      else if constexpr (F32) return launch_tb<T, C, 5, 1024, 64, 16, 4, true>(a, st);
      else return launch_tb<T, C, 5, 1024, 64, 8, 4, true>(a, st);
    }
    // depth 6: bf16 with the hand-off, 1024 lanes, 64 x 16 (110 VGPRs): 322 against 227;
    // the FP32 hand-off spills at this depth
    if (tl == 6) {
      if constexpr (B16) return launch_tb<T, C, 5, 1024, 64, 16, 6, true>(a, st);
      else return launch_tb<T, C, 5, 512, 64, 16, 6>(a, st);
    }
  }
  return hipErrorInvalidValue;
}

template <typename T, typename C>
static hipError_t tb_pass_t(const Problem &P, int tl, const void *Y, void *ZN, const void *src, long long src_idx,
                            const LaunchCfg &lc, hipStream_t st) {
  const long long sz = (long long)P.pitch * P.rows;
  TbArgs<T, C> a{};
  a.Y = (const T *)Y;
  a.ZN = (T *)ZN;
  a.coef = (const C *)P.coef;
  a.cstride = P.field_elems;
  a.sy = P.pitch;
  a.sz = sz;
  a.nx = P.nx;
  a.ny = P.ny;
  a.nz = P.nz;
  a.xcd_order = lc.xcd_order;
  a.zchunk = lc.zchunk;
  a.src = (const T *)src;
  a.src_k = src_idx < 0 ? -1 : (int)(src_idx / sz);
  a.src_off = src_idx < 0 ? 0 : (int)(src_idx % sz);
  return tb_dispatch<T, C>(P.kernel, tl, a, st);
}

thread_local int *g_fused_tile_probe = nullptr;

bool fused_tile(const Problem &P, int tl, const LaunchCfg &lc, int *tx, int *ty) {
  int t[2] = {0, 0};
  g_fused_tile_probe = t;
  const hipError_t e = tblock_pass(P, tl, nullptr, nullptr, nullptr, nullptr, nullptr, -1, lc, nullptr);
  g_fused_tile_probe = nullptr;
  if (e != hipSuccess || t[0] <= 0) return false;
  *tx = t[0];
  *ty = t[1];
  return true;
}

// One fused pass of tl levels.  Kernels 1 and 5: Y -> ZN (X, ZO unused).
// Kernel 0 (tl = 2): (X, Y) = (u^(n-1), u^n) -> (ZO, ZN) = (u^(n+1), u^(n+2)).
hipError_t tblock_pass(const Problem &P, int tl, const void *X, const void *Y, void *ZO, void *ZN, const void *src,
                       long long src_idx, const LaunchCfg &lc, hipStream_t st) {
  if (P.kernel == 0) {
    if (tl != 2) return hipErrorInvalidValue;
    return wave2_pass(P, X, Y, ZO, ZN, src, src_idx, lc, st);
  }
  if (P.kernel == 1) return jacobi_tb_pass(P, tl, Y, ZN, src, src_idx, lc, st);
  if (P.dtype == GR_F64) return tb_pass_t<double, double>(P, tl, Y, ZN, src, src_idx, lc, st);
  if (P.flags & GR_COEF_BF16) return tb_pass_t<float, __hip_bfloat16>(P, tl, Y, ZN, src, src_idx, lc, st);
  return tb_pass_t<float, float>(P, tl, Y, ZN, src, src_idx, lc, st);
}

}  // namespace gr
