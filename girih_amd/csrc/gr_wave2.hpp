// gr_wave2.hpp -- two leapfrog steps of the 25-point wave stencil per HBM pass.
//
// girih's kernel 0 (src/kernels/stencils.c:27-41) run under temporal blocking
// of depth 2: the GPU analogue of one diamond row of height 2 with the z
// wavefront of stencils_1wf.ic:19-82.  A pass reads (u^(n-1), u^n) once and
// writes (u^(n+1), u^(n+2)): 8 bytes per lattice update in FP32 against 12 for
// a single sweep.
//
// Scheme (z streaming, one barrier per plane):
//   * a workgroup of 64*BX x BY lanes owns an output tile of 64*BX x BY*RY2
//     cells and a chunk of z planes;
//   * level 1 (u^(n+1)) is computed over the tile plus an R-cell xy margin (the
//     ghost zone that makes the tile independent of its neighbours), and trails
//     the input plane by R; level 2 (u^(n+2)) is computed over the tile and
//     trails level 1 by R;
//   * lane (lx, ly) owns column lx of the tile in RY2 output rows plus 2R/BY
//     margin rows, so the output rows are the same register slots in every lane;
//     the 2R x (TY+2R) margin columns left and right of the tile belong to the
//     first few wavefronts as one more slot;
//   * each lane keeps a register queue of 2R+1 planes of u^n for its columns
//     (level 1's z neighbours; the margin-column slot keeps this queue in LDS)
//     and, for its output columns, of u^(n+1) (level 2's); the centre plane of u^n (level 1's xy neighbours, over the tile plus
//     2R) is staged in a double-buffered LDS plane, and level 1 writes u^(n+1)
//     straight into an LDS ring of R+1 planes, from which level 2 reads the xy
//     neighbours of its centre plane R steps later;
//   * every plane load goes through a buffer resource (scalar plane base, 32-bit
//     lane offset), is issued one plane ahead and is clamped into the field, so
//     loads need no predicate;
//   * the plane loop is unrolled by construction (static_for), so every queue
//     access is a fixed register.
//
// The arithmetic of each level is the single sweep's (Op<T, 0, FAST>), so a
// pass is bit-identical to two sweeps.  Boundary cells (girih's halo, never
// updated) keep their input value at level 1; the host runs this path only
// when u^(n-1) and u^n hold the same halo.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "gr_internal.hpp"
#include "gr_plane.hpp"
#include "gr_sweep.hpp"

namespace gr {


template <typename T, typename C>
struct W2Args {
  const T *P;  // u^(n-1)
  const T *V;  // u^n
  T *W;        // u^(n+1)
  T *Y;        // u^(n+2)
  const C *roc2;  // roc2 field (unless scalar), stored as C (the field type or bf16)
  T c[5];
  T csum, roc2s;
  long long sz;
  int sy;
  int nx, ny, nz;  // interior; boxes are [R, n+R)
  int tiles_x, tiles_y, nzc, zchunk, xcd_order;
  const T *src;  // src[0] is added at level 1, src[1] at level 2 (src_k < 0: no source)
  int src_k, src_off;
};

// One output tile of one pass: the body shared by the single-grid kernel and the
// batched persistent kernel (gr_batch.hip).  C is the storage type of the roc2
// field (the field type, or bf16 widened exactly to FP32 on load).
template <typename T, typename C, bool FAST, bool ROCS, bool STRM, int BX, int BY, int RY2>
struct Wave2Tile {
  static constexpr int R = 4, Q = 2 * R + 1;
  static constexpr int TX = 64 * BX, TY = BY * RY2, NT = TX * BY;
  static_assert((2 * R) % BY == 0, "the 2R margin rows must split evenly over the lane rows");
  static constexpr int RYM = 2 * R / BY;                  // margin rows per lane
  static constexpr int RY1 = RY2 + RYM;                   // level-1 rows per lane
  static constexpr int W1 = TX + 2 * R, H1 = TY + 2 * R;  // level-1 region (tile + R)
  static constexpr int W0 = TX + 4 * R, H0 = TY + 4 * R;  // level-0 region (tile + 2R)
  // LDS row pitch: a multiple of 64 words where it fits, so that the y neighbours
  // of a cell are read pairwise from one base (ds_read2st64) with no address adds
  static constexpr int LP = (sizeof(T) == 4 && BX >= 2) ? (W0 + 63) / 64 * 64 : W0;
  static constexpr int NE = 2 * R * H1;                   // level-1 cells left and right of the tile columns
  static constexpr int NEW = (NE + 63) / 64;              // wavefronts that own them
  static_assert(NEW * 64 <= NT, "margin columns need more lanes than the workgroup has");
  static constexpr int NW = R + 1;  // u^(n+1) planes in flight (written at p1, read at p2 = p1 - R)
  struct Lds {
    T lv[2][H0 * LP];    // u^n, plane p1, region 0
    T lw[NW][H1 * LP];   // u^(n+1), region 1 (same row pitch)
    T lve[Q][NEW * 64];  // z queue of the margin-column slot (its lane's own column)
  };
  __device__ static __forceinline__ void run(const W2Args<T, C> &a, int tile, int zc, Lds &sm);
};

template <typename T, typename C, bool FAST, bool ROCS, bool STRM, int BX, int BY, int RY2>
__device__ __forceinline__ void Wave2Tile<T, C, FAST, ROCS, STRM, BX, BY, RY2>::run(const W2Args<T, C> &a, int tile,
                                                                                   int zc, Lds &sm) {
  using H = Halo<R, W1, H1>;  // ring of region 0 around region 1
  constexpr int HS = (H::N + NT - 1) / NT;
  auto &lv = sm.lv;
  auto &lw = sm.lw;
  auto &lve = sm.lve;

  const int tx0 = R + (tile % a.tiles_x) * TX;  // first output cell (girih indices)
  const int ty0 = R + (tile / a.tiles_x) * TY;
  const int zs = R + zc * a.zchunk;
  const int ze = min(zs + a.zchunk, a.nz + R);
  const int nnx = a.nx + 2 * R, nny = a.ny + 2 * R, nnz = a.nz + 2 * R;
  const int tid = threadIdx.x, lx = tid % TX, ly = tid / TX;
  const bool edge_wave = tid / 64 < NEW;  // uniform per wavefront

  auto boff = [&](int gx, int gy) {
    gx = min(max(gx, 0), nnx - 1);
    gy = min(max(gy, 0), nny - 1);
    return (unsigned)(gy * a.sy + gx) * (unsigned)sizeof(T);
  };

  // main slots: column lx of the tile; rows 0..RY2-1 are output rows, the rest margin rows
  // LDS index of a region-1 cell in lw; the same cell sits at +LVO in lv
  constexpr int LVO = R * LP + R;
  unsigned moff[RY1];
  int mlw[RY1];
  bool minner[RY1], mout[RY1];
  const int gx = tx0 + lx;
#pragma unroll
  for (int j = 0; j < RY1; ++j) {
    int r1;  // row in region 1
    if (j < RY2) {
      r1 = R + ly + BY * j;
    } else {
      const int mi = ly + BY * (j - RY2);
      r1 = mi < R ? mi : TY + mi;
    }
    const int gy = ty0 - R + r1;
    moff[j] = boff(gx, gy);
    mlw[j] = r1 * LP + (lx + R);
    minner[j] = gx < a.nx + R && gy >= R && gy < a.ny + R;
    mout[j] = j < RY2 && gx < a.nx + R && gy < a.ny + R;
  }
  // margin-column slot (first NEW wavefronts)
  unsigned eoff;
  int elw;
  bool einner, evalid;
  {
    const int e = min(tid, NE - 1);
    evalid = tid < NE;
    const int side = e / (R * H1), f = e - side * (R * H1);
    const int r1 = f / R, c1 = f % R + side * (TX + R);
    const int ex = tx0 - R + c1, ey = ty0 - R + r1;
    eoff = boff(ex, ey);
    elw = r1 * LP + c1;
    einner = ex >= R && ex < a.nx + R && ey >= R && ey < a.ny + R;
  }
  // ring of region 0 (loaded from the field for each plane)
  unsigned hoff[HS];
  int hl[HS];
  bool hw[HS];
#pragma unroll
  for (int s = 0; s < HS; ++s) {
    const int c = tid + s * NT;
    int ry = 0, rx = 0;
    if (c < H::N) H::pos(c, ry, rx);
    hw[s] = c < H::N;
    hl[s] = ry * LP + rx;
    hoff[s] = boff(tx0 - 2 * R + rx, ty0 - 2 * R + ry);
  }
  const unsigned src_b = (unsigned)a.src_off * (unsigned)sizeof(T);

  const unsigned pbytes = (unsigned)a.sz * (unsigned)sizeof(T);
  auto plane = [&](const T *f, int k) { return plane_rsrc(f + (long long)min(max(k, 0), nnz - 1) * a.sz, pbytes); };
  const unsigned cbytes = (unsigned)a.sz * (unsigned)sizeof(C);
  auto cplane = [&](int k) { return plane_rsrc(a.roc2 + (long long)min(max(k, 0), nnz - 1) * a.sz, cbytes); };
  constexpr unsigned es = (unsigned)sizeof(T);  // lane offsets are in bytes of T; roc2 loads take elements
  T c[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) c[r] = a.c[r];

  // queues: u^n plane p in slot (p - zs) mod Q; u^(n+1) plane p in slot (p + R - zs) mod Q
  // (the margin-column slot, owned by a few wavefronts only, keeps its queue in LDS)
  T qv[RY1][Q], qw[RY2][Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
#pragma unroll
    for (int j = 0; j < RY1; ++j) qv[j][i] = T(0);
#pragma unroll
    for (int j = 0; j < RY2; ++j) qw[j][i] = T(0);
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {  // planes zs-2R .. zs-1
    const auto vp = plane(a.V, zs - Q + i);
#pragma unroll
    for (int j = 0; j < RY1; ++j) qv[j][i] = bload<T>(vp, moff[j]);
    if (edge_wave) lve[i][tid] = bload<T>(vp, eoff);
  }
  // operands of the first plane step (k = zs)
  T nv[RY1], nev = T(0), hv[HS], po[RY1], poe = T(0), r1v[RY1], r1e = T(0), r2v[RY2];
  auto fetch = [&](int k) {  // u^n plane k; ring, u^(n-1), roc2 of plane k-R; roc2 of plane k-2R
    const auto vp = plane(a.V, k), hp = plane(a.V, k - R), pp = plane(a.P, k - R);
#pragma unroll
    for (int j = 0; j < RY1; ++j) {
      nv[j] = bload<T>(vp, moff[j]);
      po[j] = bload<T>(pp, moff[j]);
    }
    if (edge_wave) {
      nev = bload<T>(vp, eoff);
      poe = bload<T>(pp, eoff);
    }
#pragma unroll
    for (int s = 0; s < HS; ++s) hv[s] = bload<T>(hp, hoff[s]);
    if constexpr (!ROCS) {
      const auto rp = cplane(k - R), rq = cplane(k - 2 * R);
#pragma unroll
      for (int j = 0; j < RY1; ++j) r1v[j] = bload_coef<T, C, false>(rp, moff[j] / es);
      if (edge_wave) r1e = bload_coef<T, C, false>(rp, eoff / es);
#pragma unroll
      for (int j = 0; j < RY2; ++j) r2v[j] = bload_coef<T, C, false>(rq, moff[j] / es);
    }
  };
  fetch(zs);

  const int kend = ze + 2 * R;
  int wr = 0;  // lw plane of u^(n+1) plane p1: (k - zs) mod NW; plane p2 is in (wr + 1) mod NW
  // One plane step; PH is the queue phase (k - zs) mod Q, a compile-time constant
  // so that every queue access is a fixed register.  Returns false past the end.
  auto step = [&](int kb, auto PHC) __attribute__((always_inline)) -> bool {
      constexpr int ph = decltype(PHC)::value;
      const int k = kb + ph;
      if (k >= kend) return false;
      const int p1 = k - R, p2 = k - 2 * R;  // planes of level 1 and level 2 in this step
      const int s1 = (ph + Q - R) % Q;  // queue slot of u^n plane p1
      T *LV = lv[k & 1];
      T *LW1 = lw[wr];                          // u^(n+1) plane p1 (written below)
      const T *LW2 = lw[wr == NW - 1 ? 0 : wr + 1];  // u^(n+1) plane p2
      wr = wr == NW - 1 ? 0 : wr + 1;
      // u^n plane k enters the queue; LDS gets u^n plane p1 (queue, ring)
#pragma unroll
      for (int j = 0; j < RY1; ++j) qv[j][ph] = nv[j];
      if (edge_wave) lve[ph][tid] = nev;
#pragma unroll
      for (int j = 0; j < RY1; ++j) LV[mlw[j] + LVO] = qv[j][s1];
#pragma unroll
      for (int s = 0; s < HS; ++s)
        if (hw[s]) LV[hl[s]] = hv[s];
      if (edge_wave && evalid) LV[elw + LVO] = lve[s1][tid];
      T uo1[RY1], uoe = poe, rc1[RY1], rce = ROCS ? a.roc2s : r1e, rc2[RY2];
#pragma unroll
      for (int j = 0; j < RY1; ++j) {
        uo1[j] = po[j];
        rc1[j] = ROCS ? a.roc2s : r1v[j];
      }
#pragma unroll
      for (int j = 0; j < RY2; ++j) rc2[j] = ROCS ? a.roc2s : r2v[j];
      fetch(k + 1);
      __syncthreads();

      // ---- level 1: u^(n+1) at plane p1 over the tile plus R, into the LDS ring
      // (plane p1's ring entry was last read as plane p2 in the previous step)
      const bool zin1 = p1 >= R && p1 < a.nz + R;
      const bool src1 = a.src_k >= 0 && p1 == a.src_k;  // src_k = -1: no source
      auto level1 = [&](auto q, unsigned off, int lw1, bool inner, T uo, T roc) {
        // neighbours at non-negative offsets from the cell's (-R, -R) corner, so that
        // every LDS read takes its offset as an immediate
        const T *Lc = LV + lw1;
        auto xn = [&](int d) { return Lc[LVO + d]; };
        auto yn = [&](int d) { return Lc[(R + d) * LP + R]; };
        T zq[Q];
#pragma unroll
        for (int d = -R; d <= R; ++d) zq[R + d] = q((ph + Q - R + d) % Q);
        const T v0 = zq[R];
        T w = Op<T, 0, FAST>::update(v0, uo, roc, Op<T, 0, FAST>::template lap<kFastND>(c, a.csum, v0, zq, xn, yn));
        if (src1 && off == src_b) w = w + a.src[0];
        return (inner && zin1) ? w : v0;  // girih's halo keeps its value
      };
#pragma unroll
      for (int j = 0; j < RY1; ++j) {
        const T w = level1([&](int i) { return qv[j][i]; }, moff[j], mlw[j], minner[j], uo1[j], rc1[j]);
        LW1[mlw[j]] = w;
        if (j < RY2) qw[j < RY2 ? j : 0][ph] = w;
      }
      if (edge_wave && evalid) LW1[elw] = level1([&](int i) { return lve[i][tid]; }, eoff, elw, einner, uoe, rce);

      // ---- level 2: u^(n+2) at plane p2 over the tile
      if (k >= zs + 2 * R) {
        const bool src2 = a.src_k >= 0 && p2 == a.src_k;
        const auto wp = plane_rsrc(a.W + (long long)p2 * a.sz, pbytes);
        const auto yp = plane_rsrc(a.Y + (long long)p2 * a.sz, pbytes);
#pragma unroll
        for (int j = 0; j < RY2; ++j) {
          const T *Lc = LW2 + (mlw[j] - LVO);  // (-R, -R) corner, as in level 1
          auto xn = [&](int d) { return Lc[LVO + d]; };
          auto yn = [&](int d) { return Lc[(R + d) * LP + R]; };
          T zq[Q];
#pragma unroll
          for (int d = -R; d <= R; ++d) zq[R + d] = qw[j][(ph + Q - R + d) % Q];
          const T w0 = zq[R];
          const T uo = qv[j][(ph + 1) % Q];  // u^n plane p2
          T y = Op<T, 0, FAST>::update(w0, uo, rc2[j],
                                       Op<T, 0, FAST>::template lap<kFastND>(c, a.csum, w0, zq, xn, yn));
          if (src2 && moff[j] == src_b) y = y + a.src[1];
          if (mout[j]) {
            bstore<STRM>(w0, wp, moff[j]);
            bstore<STRM>(y, yp, moff[j]);
          }
        }
      }
      return true;
  };
  for (int kb = zs; kb < kend; kb += Q) static_for<Q>([&](auto PHC) __attribute__((always_inline)) {
    return step(kb, PHC);
  });
}

}  // namespace gr
