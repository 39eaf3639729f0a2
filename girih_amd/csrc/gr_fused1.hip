// gr_fused1.hip -- TL steps of the 7-point constant-coefficient stencil per HBM pass.
//
// girih's kernel 1 (src/kernels/stencils.c:73-78) under temporal blocking of
// depth TL: the GPU analogue of a diamond of height TL with the z wavefront of
// stencils_1wf.ic:19-82.  A pass reads the field once and writes it once: 2
// words per cell for TL updates.
//
// Scheme (z streaming, one barrier per plane step):
//   * a workgroup owns an output tile of TX x TY cells and a chunk of z planes;
//     every level is computed over the tile plus a margin of TL-1 cells (the
//     ghost zone of the last level that makes the tile independent);
//   * lanes own cells of that region in row-major order (NCOL slots each) and
//     keep, per level, a register queue of the last 3 planes of their cells;
//   * in plane step k, level t computes plane k-t: its z neighbours are level
//     t-1's planes k-t-1 .. k-t+1 (the same lane's queue, the newest computed in
//     this step), its xy neighbours come from level t-1's plane k-t in LDS,
//     written one step earlier (levels >= 1) or at the start of this step (the
//     input).  Every level writes its plane into a double-buffered LDS slot, so
//     one barrier per step orders all levels;
//   * the plane loop is unrolled by 6 (queue slot k mod 3, LDS slot k mod 2 are
//     compile-time constants); plane loads go through buffer resources.
//
// The arithmetic is the single sweep's (girih's operand order, no contraction),
// so a pass is bit-identical to TL sweeps.  Boundary cells (girih's halo) keep
// their input value at every level; the host runs this path only when U1 and U2
// hold the same halo.
#include <hip/hip_runtime.h>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"
#include "gr_plane.hpp"
#include "gr_sweep.hpp"

namespace gr {

template <typename T>
struct J1Args {
  const T *Y;  // input level
  T *Z;        // level TL
  T c0, c1;
  long long sz;
  int sy;
  int nx, ny, nz;  // interior; boxes are [1, n+1)
  int tiles_x, tiles_y, nzc, zchunk, xcd_order;
  const T *src;  // level t adds src[t-1] (src_k < 0: no source)
  int src_k, src_off;
};

template <typename T, int TL, int TX, int TY, int NT, bool STRM>
__global__ __launch_bounds__(NT) void jacobi_tb_kernel(const J1Args<T> a) {
  constexpr int M1 = TL - 1;                       // margin of the computed region
  constexpr int W1 = TX + 2 * M1, H1 = TY + 2 * M1;  // computed region
  constexpr int W0 = W1 + 2, H0 = H1 + 2;          // plus the input's ring; LDS planes
  constexpr int NCOL = (W1 * H1 + NT - 1) / NT;
  using H = Halo<1, W1, H1>;
  constexpr int HS = (H::N + NT - 1) / NT;
  constexpr int UN = 6;  // plane steps per unrolled block: lcm of the queue (3) and LDS (2) rings
  __shared__ T lds[TL][2][H0 * W0];  // level t's latest planes (t = 0: the input)

  const int ntile = a.tiles_x * a.tiles_y;
  const int lb = tile_order(blockIdx.x, ntile * a.nzc, ntile, a.xcd_order);
  const int tile = lb % ntile;
  const int tx0 = 1 + (tile % a.tiles_x) * TX;  // first output cell (girih indices)
  const int ty0 = 1 + (tile / a.tiles_x) * TY;
  const int zs = 1 + (lb / ntile) * a.zchunk;
  const int ze = min(zs + a.zchunk, a.nz + 1);
  const int nnx = a.nx + 2, nny = a.ny + 2, nnz = a.nz + 2;
  const int tid = threadIdx.x;

  auto boff = [&](int gx, int gy) {
    gx = min(max(gx, 0), nnx - 1);
    gy = min(max(gy, 0), nny - 1);
    return (unsigned)(gy * a.sy + gx) * (unsigned)sizeof(T);
  };
  unsigned off[NCOL];
  int lo[NCOL];  // LDS index of the cell's (-1, -1) neighbour: every read has a non-negative immediate offset
  bool valid[NCOL], inner[NCOL], outc[NCOL];
#pragma unroll
  for (int j = 0; j < NCOL; ++j) {
    const int c = min(tid + j * NT, W1 * H1 - 1);
    valid[j] = tid + j * NT < W1 * H1;
    const int cx = c % W1, cy = c / W1;
    const int gx = tx0 - M1 + cx, gy = ty0 - M1 + cy;
    off[j] = boff(gx, gy);
    lo[j] = cy * W0 + cx;
    inner[j] = gx >= 1 && gx < a.nx + 1 && gy >= 1 && gy < a.ny + 1;
    outc[j] = valid[j] && cx >= M1 && cx < M1 + TX && cy >= M1 && cy < M1 + TY && gx < a.nx + 1 && gy < a.ny + 1;
  }
  unsigned hoff[HS];
  int hl[HS];
  bool hw[HS];
#pragma unroll
  for (int s = 0; s < HS; ++s) {
    const int c = tid + s * NT;
    int ry = 0, rx = 0;
    if (c < H::N) H::pos(c, ry, rx);
    hw[s] = c < H::N;
    hl[s] = ry * W0 + rx;
    hoff[s] = boff(tx0 - M1 - 1 + rx, ty0 - M1 - 1 + ry);
  }
  const unsigned src_b = (unsigned)a.src_off * (unsigned)sizeof(T);
  const unsigned pbytes = (unsigned)a.sz * (unsigned)sizeof(T);
  auto plane = [&](const T *f, int k) { return plane_rsrc(f + (long long)min(max(k, 0), nnz - 1) * a.sz, pbytes); };
  const T c0 = a.c0, c1 = a.c1;

  // queues: level t's plane p in slot (p + t - kfirst) mod 3, i.e. the step that computed it
  T q[TL][NCOL][3];
#pragma unroll
  for (int t = 0; t < TL; ++t)
#pragma unroll
    for (int j = 0; j < NCOL; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) q[t][j][i] = T(0);

  const int kfirst = zs - TL, kend = ze + TL;  // step k: the input's plane k enters; level t computes k - t
  T nv[NCOL], hv[HS];
  auto fetch = [&](int k) {  // input plane k (own cells) and the ring of input plane k - 1
    const auto vp = plane(a.Y, k), hp = plane(a.Y, k - 1);
#pragma unroll
    for (int j = 0; j < NCOL; ++j) nv[j] = bload<T>(vp, off[j]);
#pragma unroll
    for (int s = 0; s < HS; ++s) hv[s] = bload<T>(hp, hoff[s]);
  };
  fetch(kfirst);

  auto step = [&](int kb, auto PHC) __attribute__((always_inline)) -> bool {
    constexpr int ph = decltype(PHC)::value;
    constexpr int qs = ph % 3, qm1 = (ph + 2) % 3, qm2 = (ph + 1) % 3;  // slots of steps k, k-1, k-2
    constexpr int bs = ph % 2, bm1 = (ph + 1) % 2;                      // LDS slots of steps k, k-1
    const int k = kb + ph;
    if (k >= kend) return false;
    // the input's plane k enters its queue; its plane k-1 (level 1's centre plane) goes to LDS
#pragma unroll
    for (int j = 0; j < NCOL; ++j) q[0][j][qs] = nv[j];
    {
      T *L0 = lds[0][bs];
#pragma unroll
      for (int j = 0; j < NCOL; ++j)
        if (valid[j]) L0[lo[j] + W0 + 1] = q[0][j][qm1];
#pragma unroll
      for (int s = 0; s < HS; ++s)
        if (hw[s]) L0[hl[s]] = hv[s];
    }
    fetch(k + 1);
    __syncthreads();

#pragma unroll
    for (int t = 1; t <= TL; ++t) {
      const int p = k - t;
      const bool zin = p >= 1 && p < a.nz + 1;
      // (p runs below plane 0 for the first chunk: src_k = -1 means "no source" and must not match it)
      const bool srcp = a.src_k >= 0 && p == a.src_k;
      const T *Lp = lds[t - 1][t == 1 ? bs : bm1];  // level t-1's plane p
#pragma unroll
      for (int j = 0; j < NCOL; ++j) {
        const T *Lc = Lp + lo[j];
        const T v0 = q[t - 1][j][qm1];                        // plane p
        const T zm = q[t - 1][j][qm2], zp = q[t - 1][j][qs];  // planes p-1, p+1
        // stencils.c:73-78, operand order kept
        T out = c0 * v0 + c1 * (Lc[W0 + 2] + Lc[W0]) + c1 * (Lc[2 * W0 + 1] + Lc[1]) + c1 * (zp + zm);
        if (srcp && off[j] == src_b) out = out + a.src[t - 1];
        const T val = (inner[j] && zin) ? out : v0;  // girih's halo keeps its value
        if (t < TL) {
          q[t < TL ? t : 0][j][qs] = val;
          if (valid[j]) lds[t < TL ? t : 0][bs][lo[j] + W0 + 1] = val;
        } else if (outc[j] && p >= zs && p < ze) {
          bstore<STRM>(val, plane(a.Z, p), off[j]);
        }
      }
    }
    return true;
  };
  for (int kb = kfirst; kb < kend; kb += UN) static_for<UN>([&](auto PHC) __attribute__((always_inline)) {
    return step(kb, PHC);
  });
}

template <typename T, int TL, int TX, int TY, int NT, bool STRM>
static hipError_t launch_j1(J1Args<T> a, hipStream_t st) {
  if (g_fused_tile_probe) {
    g_fused_tile_probe[0] = TX;
    g_fused_tile_probe[1] = TY;
    return hipSuccess;
  }
  a.tiles_x = (a.nx + TX - 1) / TX;
  a.tiles_y = (a.ny + TY - 1) / TY;
  a.zchunk = (a.zchunk <= 0 || a.zchunk > a.nz) ? a.nz : a.zchunk;
  a.nzc = (a.nz + a.zchunk - 1) / a.zchunk;
  hipLaunchKernelGGL((jacobi_tb_kernel<T, TL, TX, TY, NT, STRM>), dim3(a.tiles_x * a.tiles_y * a.nzc), dim3(NT), 0,
                     st, a);
  return hipGetLastError();
}

// Tiles: TL*2 LDS planes of (TX + 2TL) x (TY + 2TL) words must fit the 160 KB (the
// compiler rejects a kernel whose LDS exceeds it).
template <typename T, bool STRM>
static hipError_t j1_dispatch(int tl, const J1Args<T> &a, hipStream_t st) {
  if constexpr (sizeof(T) == 4) {
    if (tl == 2) return launch_j1<T, 2, 128, 32, 1024, STRM>(a, st);
    if (tl == 3) return launch_j1<T, 3, 128, 16, 1024, STRM>(a, st);
    if (tl == 4) return launch_j1<T, 4, 128, 16, 1024, STRM>(a, st);
    if (tl == 5) return launch_j1<T, 5, 128, 16, 1024, STRM>(a, st);
    if (tl == 6) return launch_j1<T, 6, 64, 16, 1024, STRM>(a, st);
  } else {
    if (tl == 2) return launch_j1<T, 2, 128, 16, 1024, STRM>(a, st);
    if (tl == 3) return launch_j1<T, 3, 64, 16, 512, STRM>(a, st);  // 640 against 622 GLUP/s at 128 x 16
    if (tl == 4) return launch_j1<T, 4, 64, 16, 1024, STRM>(a, st);
    if (tl == 5) return launch_j1<T, 5, 64, 16, 1024, STRM>(a, st);
    if (tl == 6) return launch_j1<T, 6, 64, 8, 512, STRM>(a, st);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t j1_pass_t(const Problem &P, int tl, const void *Y, void *Z, const void *src, long long src_idx,
                            const LaunchCfg &lc, hipStream_t st) {
  const long long sz = (long long)P.pitch * P.rows;
  J1Args<T> a{};
  a.Y = (const T *)Y;
  a.Z = (T *)Z;
  a.c0 = (T)P.c[0];
  a.c1 = (T)P.c[1];
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
  if (streaming_sweep(P)) return j1_dispatch<T, true>(tl, a, st);
  return j1_dispatch<T, false>(tl, a, st);
}

// One pass of tl kernel-1 steps: Y -> Z.
hipError_t jacobi_tb_pass(const Problem &P, int tl, const void *Y, void *Z, const void *src, long long src_idx,
                          const LaunchCfg &lc, hipStream_t st) {
  if (P.kernel != 1) return hipErrorInvalidValue;
  if (P.dtype == GR_F64) return j1_pass_t<double>(P, tl, Y, Z, src, src_idx, lc, st);
  return j1_pass_t<float>(P, tl, Y, Z, src, src_idx, lc, st);
}

}  // namespace gr
