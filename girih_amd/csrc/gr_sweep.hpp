// gr_sweep.hpp -- single-sweep star-stencil kernels for gfx950 (CDNA4).
//
// One launch applies one time step u <- f(v) over a box of a 3-D field
// (girih's KERNEL_SIG sweep, src/kernels/stencils_spt_blk.ic:19-48, with the
// FUNC_BODY operators of src/kernels/stencils.c:27-216).
//
// Scheme (2.5-D z streaming):
//   * a workgroup of TX = 64*BX by BY lanes owns an output tile of TX (x) by
//     BY*RY (y) cells and marches through a chunk of z planes (wider tiles read
//     fewer partial halo lines per interior line);
//   * each lane keeps, for its RY cells, a register queue of the 2R+1 planes
//     k-R .. k+R of v (the z neighbours never touch LDS);
//   * the current plane k of the tile plus its R-wide xy halo is staged in LDS
//     (double buffered, one barrier per plane); x and y neighbours come from LDS;
//   * every global load (next queue plane, next halo ring, u^(n-1), coefficient
//     arrays) is issued one plane ahead, so HBM latency overlaps the arithmetic;
//   * x is the 64-lane wavefront axis: every load/store of a row is one
//     coalesced 256 B (FP32) / 512 B (FP64) transaction;
//   * workgroups may be renumbered (tile_order) so that each XCD processes a
//     contiguous run of tiles: tiles that share halos run on the same XCD and hit
//     its L2.
//
// Arithmetic modes:
//   * reference order (default): the association and operand order of girih's
//     FUNC_BODY, compiled with -ffp-contract=off, so FP64 and FP32 results are
//     bit-identical to girih's serial reference;
//   * FAST (FP32, kernel 0): per-radius neighbour sums (radii 1-2 as
//     differences from the centre), FMA accumulation and the increment form
//     u^n + ((u^n - u^(n-1)) + roc2*lap) of the leapfrog update.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <type_traits>

#include "gr_plane.hpp"

#ifndef GR_FAST_ND
#define GR_FAST_ND 2
#endif

#ifndef GR_STREAMING_HINTS
#define GR_STREAMING_HINTS 1
#endif

namespace gr {

// FAST kernel 0: radii 1..kFastND use the difference form (see Op<T, 0, true>).
constexpr int kFastND = GR_FAST_ND;

// Centre weight of the FAST Laplacian, c0 + 6*sum_{r <= kFastND} c_r, rounded once from FP64.
inline float fast_csum(const double *c) {
  double s = c[0];
  for (int r = 1; r <= kFastND; ++r) s += 6.0 * c[r];
  return (float)s;
}

__host__ __device__ constexpr int kradius(int kid) { return (kid == 0 || kid == 4) ? 4 : 1; }
__host__ __device__ constexpr int kncoef(int kid) {
  return kid == 2 ? 2 : kid == 3 ? 4 : kid == 4 ? 13 : kid == 5 ? 7 : 0;
}

// Field geometry: element (i, j, k) of a field lives at base[k*sz + j*sy + i]
// where base already includes the x offset of logical x = 0.
struct Geom {
  long long sy, sz;
  int xb, yb, zb, xe, ye, ze;  // output box
};

template <typename T, typename C>
struct SweepArgs {
  T *u;                // output (kernel 0: holds u^(n-1) on entry unless uold is set)
  const T *v;          // input field
  const T *uold;       // kernel 0: u^(n-1) when it lives in another field than u (null = u)
  const C *roc2;       // kernel 0 roc2 field (unless scalar)
  const C *coef;       // coefficient arrays, array m at coef + m*cstride
  long long cstride;
  T c[5];              // constant coefficients
  T csum;              // FAST kernel 0 centre weight (fast_csum)
  T roc2s;             // scalar roc2
  Geom g;
  int tiles_x, tiles_y, nzc, zchunk;
  int xcd_order;       // workgroup -> tile order (tile_order): 0 plain, 1 XCD runs, 2 XCD bands per z layer
  const T *src;        // point source value, added at plane src_k, in-plane offset src_off
  int src_k;           // -1 = no source
  int src_off;
};

// Coefficient and roc2 fields are read once per sweep: with NT, loads carry the
// non-temporal hint (see ldb_stream).  bf16 values widen exactly (the bits move to
// the top half).
template <bool NT, typename X>
__device__ __forceinline__ X ld_once(const X *p) {
  if constexpr (NT && GR_STREAMING_HINTS) return __builtin_nontemporal_load(p);
  return *p;
}
template <typename T, typename C, bool NT>
__device__ __forceinline__ T ld_coef(const C *p) {
  if constexpr (std::is_same<C, __hip_bfloat16>::value) {
    const unsigned short b = ld_once<NT>(reinterpret_cast<const unsigned short *>(p));
    return static_cast<T>(__uint_as_float((unsigned)b << 16));
  } else {
    return static_cast<T>(ld_once<NT>(p));
  }
}

// Loads/stores through a uniform base and a 32-bit byte offset: the compiler then
// uses the SGPR-base + VGPR-offset global addressing mode (no per-access 64-bit
// address arithmetic in the vector ALU).
template <typename X>
__device__ __forceinline__ X ldb(const void *base, unsigned boff) {
  return *reinterpret_cast<const X *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename X>
__device__ __forceinline__ void stb(void *base, unsigned boff, X v) {
  *reinterpret_cast<X *>(reinterpret_cast<char *>(base) + boff) = v;
}
// Streaming variants for data touched once per sweep (the output plane and the
// u^(n-1) operand).  When a sweep's working set is far larger than the caches
// (NT), non-temporal hints keep these from displacing the halo lines that
// neighbouring tiles still read from L2; a small grid that stays in the cache
// between steps loads and stores normally.  The choice is a template parameter: a runtime
// choice between the two forms of one access lets the compiler merge them and drop
// the hint.
template <typename X, bool NT>
__device__ __forceinline__ X ldb_stream(const void *base, unsigned boff) {
  if constexpr (NT && GR_STREAMING_HINTS)
    return __builtin_nontemporal_load(reinterpret_cast<const X *>(reinterpret_cast<const char *>(base) + boff));
  return ldb<X>(base, boff);
}
template <typename X, bool NT>
__device__ __forceinline__ void stb_stream(void *base, unsigned boff, X v) {
  if constexpr (NT && GR_STREAMING_HINTS)
    __builtin_nontemporal_store(v, reinterpret_cast<X *>(reinterpret_cast<char *>(base) + boff));
  else
    stb<X>(base, boff, v);
}

// Map a linear workgroup id to a logical tile so that each of the 8 XCDs (which
// receive workgroups round-robin) owns one contiguous range of tiles.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7;
  const int xcd = b & 7, loc = b >> 3;
  return (xcd < rem) ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// Workgroup -> logical item (tile + z chunk, tile fastest) for `order`:
//   0: plain (the XCDs take tiles round-robin: with 8 tiles across, a tile column each);
//   1: each XCD owns a contiguous run of the whole grid's items (its own z chunks);
//   2: each XCD owns a contiguous run of the tiles of every z layer (`layer` items): a
//      band of whole tile rows, while all XCDs stay on the same z chunk.
// Every order is a bijection of [0, nb); nb is a multiple of layer.
__device__ __forceinline__ int tile_order(int b, int nb, int layer, int order) {
  if (order == 1) return xcd_remap(b, nb);
  if (order == 2) {
    const int z = b / layer;
    return z * layer + xcd_remap(b - z * layer, layer);
  }
  return b;
}

template <int R, int TX, int TY>
struct Halo {
  static constexpr int LX = TX + 2 * R;
  static constexpr int LY = TY + 2 * R;
  static constexpr int N = 2 * R * LX + 2 * R * TY;
  // halo slot c -> (ly, lx) inside the LDS tile
  __device__ static __forceinline__ void pos(int c, int &ly, int &lx) {
    if (c < R * LX) {
      ly = c / LX;
      lx = c - ly * LX;
    } else if (c < 2 * R * LX) {
      const int d = c - R * LX;
      ly = R + TY + d / LX;
      lx = d % LX;
    } else {
      const int d = c - 2 * R * LX;
      const int side = d / (R * TY);
      const int e = d - side * R * TY;
      ly = R + e / R;
      lx = (e % R) + side * (TX + R);
    }
  }
};

template <typename T, int KID, bool FAST>
struct Op;

// ---- kernel 0: 25-point star, second order in time (stencils.c:27-41) ----
template <typename T>
struct Op<T, 0, false> {
  template <int ND, typename X, typename Y>
  __device__ static __forceinline__ T lap(const T *c, T, T v0, const T *q, X xn, Y yn) {
    // coef[0]*v + coef[r]*(x pair) + coef[r]*(y pair) + coef[r]*(z pair), r = 1..4, left to right
    T l = c[0] * v0;
#pragma unroll
    for (int r = 1; r <= 4; ++r) {
      l = l + c[r] * (xn(r) + xn(-r));
      l = l + c[r] * (yn(r) + yn(-r));
      l = l + c[r] * (q[4 + r] + q[4 - r]);
    }
    return l;
  }
  __device__ static __forceinline__ T update(T v0, T uo, T roc2, T l) {
    return (T(2) * v0 - uo) + roc2 * l;
  }
};

// FAST kernel 0 (FP32).  Radii r <= ND use the difference form
//   c_r * sum_axes((a - v) + (b - v)),
// radii r > ND the plain pair sums c_r * sum_axes(a + b); the centre weight is
// csum = c0 + 6*sum_{r<=ND} c_r, rounded once from FP64 on the host.  The
// differences of a smooth field are small and nearly exact, which removes the
// cancellation of c0*v against the large-weight neighbour sums that dominates the
// FP32 error of the plain form; the small far-radius weights (|c3|, |c4| << |c1|)
// keep the residual cancellation small at 12 fewer operations than ND = 4.
template <typename T>
struct Op<T, 0, true> {
  template <int ND, typename X, typename Y>
  __device__ static __forceinline__ T lap(const T *c, T csum, T v0, const T *q, X xn, Y yn) {
    T l = csum * v0;
#pragma unroll
    for (int r = 4; r >= 1; --r) {
      T d;
      if (r <= ND)
        d = (((xn(r) - v0) + (xn(-r) - v0)) + ((yn(r) - v0) + (yn(-r) - v0))) +
            ((q[4 + r] - v0) + (q[4 - r] - v0));
      else
        d = ((xn(r) + xn(-r)) + (yn(r) + yn(-r))) + (q[4 + r] + q[4 - r]);
      l = __builtin_fmaf(c[r], d, l);
    }
    return l;
  }
  // u^(n+1) = u^n + ((u^n - u^(n-1)) + roc2*lap): the increment form of the leapfrog.
  // u^n - u^(n-1) of a smoothly evolving field is exact (Sterbenz), the small
  // increment takes the Laplacian term with one rounding, and only the final add
  // rounds at the magnitude of u -- instead of 2u^n - u^(n-1) losing the low bits
  // of the increment.
  __device__ static __forceinline__ T update(T v0, T uo, T roc2, T l) {
    return v0 + __builtin_fmaf(roc2, l, v0 - uo);
  }
};

// Output tile per workgroup: 64*BX x BY lanes, RY rows per lane (64*BX x BY*RY cells).
template <typename T, int KID>
struct Tile {
  static constexpr int BY = (KID == 4) ? 8 : 4;
  static constexpr int RY = (KID == 4) ? 2 : 4;
};

// BUF: every plane access through a buffer resource (scalar plane base, 32-bit lane
// offset) instead of global loads; measured per case, see the launch code.
// PF (kernel 0): the u^(n-1) and roc2 operands of plane k+1 are loaded while plane k
// is computed, so their latency overlaps a whole plane instead of the LDS exchange.
template <typename T, typename C, int KID, int BY, int RY, bool FAST, bool ROCS, bool RING, int ND, int BX = 1,
          bool STRM = false, bool BUF = false, bool PF = false>
struct SweepTile;

// MINW > 0 asks the compiler for at least MINW waves per SIMD (caps VGPRs at 512/MINW).
// RING: the LDS holds a ring of R+2 tile planes and the halo ring of a plane is
// fetched in the same step as the interior of that plane (R+1 planes ahead), when
// the neighbouring tiles fetch it too and it is still in L2; otherwise the LDS is
// double buffered and the halo is fetched one plane ahead.
template <typename T, typename C, int KID, int BY, int RY, bool FAST, bool ROCS, int MINW, bool RING, int ND = kFastND,
          int BX = 1, bool STRM = false, bool BUF = false, bool PF = false>
__global__ __launch_bounds__(64 * BX * BY) __attribute__((amdgpu_waves_per_eu(MINW > 0 ? MINW : 1)))
void sweep_kernel(const SweepArgs<T, C> a) {
  using ST = SweepTile<T, C, KID, BY, RY, FAST, ROCS, RING, ND, BX, STRM, BUF, PF>;
  __shared__ typename ST::Lds lds;
  const int nb = a.tiles_x * a.tiles_y * a.nzc;
  const int lb = tile_order(blockIdx.x, nb, a.tiles_x * a.tiles_y, a.xcd_order);
  const int tile_x = lb % a.tiles_x;
  const int tile_y = (lb / a.tiles_x) % a.tiles_y;
  const int zc = lb / (a.tiles_x * a.tiles_y);
  ST::run(a, tile_x, tile_y, zc, lds);
}

// One tile of one sweep: the z-streaming body shared by the single-grid kernel
// and the batched persistent kernel.
template <typename T, typename C, int KID, int BY, int RY, bool FAST, bool ROCS, bool RING, int ND, int BX, bool STRM,
          bool BUF, bool PF>
struct SweepTile {
  static constexpr int R = kradius(KID);
  static constexpr int TX = 64 * BX, TY = BY * RY;
  static constexpr int LX = Halo<R, TX, TY>::LX, LY = Halo<R, TX, TY>::LY;
  static constexpr int S = RING ? R + 2 : 2;  // LDS tile planes
  typedef T Lds[S][LY][LX];
  __device__ static __forceinline__ void run(const SweepArgs<T, C> &a, int tile_x, int tile_y, int zc, Lds &lds);
};

template <typename T, typename C, int KID, int BY, int RY, bool FAST, bool ROCS, bool RING, int ND, int BX, bool STRM,
          bool BUF, bool PF>
__device__ __forceinline__ void SweepTile<T, C, KID, BY, RY, FAST, ROCS, RING, ND, BX, STRM, BUF, PF>::run(const SweepArgs<T, C> &a,
                                                                                   int tile_x, int tile_y, int zc,
                                                                                   Lds &lds) {
  constexpr int Q = 2 * R + 1;
  constexpr int NT = TX * BY;
  constexpr int NC = kncoef(KID);
  using H = Halo<R, TX, TY>;
  constexpr int HS = (H::N + NT - 1) / NT;  // halo slots per lane

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * TX + tx;
  const Geom &g = a.g;
  const int x0 = g.xb + tile_x * TX;
  const int y0 = g.yb + tile_y * TY;
  const int z0 = g.zb + zc * a.zchunk;
  const int z1 = min(z0 + a.zchunk, g.ze);

  const int gx = x0 + tx;
  const bool xout = gx < g.xe;
  // Every load address is clamped into the part of the field a tile may read
  // ([xb-R, xe+R) x [yb-R, ye+R)), so loads need no predicate or branch; lanes
  // outside the output box compute garbage they never store.
  const int gxc = min(gx, g.xe + R - 1);

  unsigned off[RY];  // in-plane element offsets (fit 32 bits); plane bases are uniform
  bool yout[RY];
#pragma unroll
  for (int m = 0; m < RY; ++m) {
    const int gy = y0 + ty * RY + m;
    off[m] = (unsigned)(min(gy, g.ye + R - 1) * (int)g.sy + gxc) * (unsigned)sizeof(T);  // bytes
    yout[m] = xout && (gy < g.ye);
  }
  int hl[HS];
  unsigned hoff[HS];
  bool hw[HS];
#pragma unroll
  for (int s = 0; s < HS; ++s) {
    const int c = tid + s * NT;
    int ly = 0, lx = 0;
    if (c < H::N) H::pos(c, ly, lx);
    const int hx = min(x0 - R + lx, g.xe + R - 1), hy = min(y0 - R + ly, g.ye + R - 1);
    hl[s] = ly * LX + lx;
    hoff[s] = (unsigned)(hy * (int)g.sy + hx) * (unsigned)sizeof(T);  // bytes
    hw[s] = c < H::N;
  }
  auto plane = [&](const T *base, int k) { return base + (long long)k * g.sz; };
  const int zlast = g.ze + R - 1;  // last plane any load may touch
  // plane accesses: buffer resources (BUF) or SGPR-base global loads
  const unsigned pbytes = (unsigned)g.sz * (unsigned)sizeof(T);
  auto pl = [&](const T *base, int k) {
    if constexpr (BUF)
      return plane_rsrc(base + (long long)k * g.sz, pbytes);
    else
      return plane(base, k);
  };
  auto ld = [&](auto p, unsigned o) -> T {
    if constexpr (BUF)
      return bload<T>(p, o);
    else
      return ldb<T>(p, o);
  };
  auto ld_strm = [&](auto p, unsigned o) -> T {
    if constexpr (BUF)
      return bload<T, STRM>(p, o);
    else
      return ldb_stream<T, STRM>(p, o);
  };

  // z queue of the tile interior: slot i holds plane k-R+((i - p) mod Q) at phase p;
  // the z loop is unrolled by Q so the rotation is static register renaming.
  T q[RY][Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const auto vp = pl(a.v, z0 - R + i);
#pragma unroll
    for (int m = 0; m < RY; ++m) q[m][i] = ld(vp, off[m]);
  }
  T hv[HS];
  if constexpr (RING) {
    for (int j = 0; j < R; ++j) {  // halo rings of planes z0 .. z0+R-1 straight into their slots
      const auto vp = pl(a.v, z0 + j);
      T *Ls = &lds[(z0 + j) % S][0][0];
#pragma unroll
      for (int s = 0; s < HS; ++s)
        if (hw[s]) Ls[hl[s]] = ld(vp, hoff[s]);
    }
  }
  {
    const auto vp = pl(a.v, min(RING ? z0 + R : z0, zlast));
#pragma unroll
    for (int s = 0; s < HS; ++s) hv[s] = ld(vp, hoff[s]);
  }
  T c[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) c[r] = a.c[r];

  // kernel 0's own-cell operands of plane k: u^(n-1) and, with a roc2 field, roc2.  The
  // roc2 plane is read through a buffer resource (scalar plane base, 32-bit lane
  // offset): 11% faster for kernel 0 with a bf16 roc2 field at 1024^3.
  auto load_u0 = [&](int k, T (&u_)[RY], T (&r_)[RY]) {
    if constexpr (KID == 0) {
      const auto up = pl(a.uold ? a.uold : a.u, k);
      const unsigned cbytes = (unsigned)g.sz * (unsigned)sizeof(C);
#pragma unroll
      for (int m = 0; m < RY; ++m) {
        u_[m] = ld_strm(up, off[m]);
        if constexpr (!ROCS)
          r_[m] = bload_coef<T, C, STRM>(plane_rsrc(a.roc2 + (long long)k * g.sz, cbytes), off[m] / (unsigned)sizeof(T));
        else
          r_[m] = T(0);
      }
    }
  };
  T uon[RY], r2n[RY];  // PF: operands of the next plane
  if constexpr (PF) load_u0(z0, uon, r2n);

  for (int kb = z0; kb < z1; kb += Q) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const int k = kb + p;
      if (k >= z1) break;
      const int buf = RING ? k % S : (k - z0) & 1;
      T(*L)[LX] = lds[buf];
#pragma unroll
      for (int m = 0; m < RY; ++m) L[R + ty * RY + m][R + tx] = q[m][(p + R) % Q];
      {
        T *Lh = RING ? &lds[(k + R) % S][0][0] : &L[0][0];  // halo of plane k+R (ring) or k
#pragma unroll
        for (int s = 0; s < HS; ++s)
          if (hw[s]) Lh[hl[s]] = hv[s];
      }

      // operands of plane k (arrive while the LDS tile is written and synchronized,
      // or, with PF, while plane k-1 was computed)
      T uo[RY], r2[RY];
      T cf[RY][NC > 0 ? NC : 1];
      if constexpr (PF) {
#pragma unroll
        for (int m = 0; m < RY; ++m) {
          uo[m] = uon[m];
          r2[m] = r2n[m];
        }
        // plane k+1, kept inside this workgroup's chunk: the next chunk's first plane is
        // being written by another workgroup (the value loaded past the end is unused)
        load_u0(min(k + 1, z1 - 1), uon, r2n);
      } else {
        load_u0(k, uo, r2);
      }
      if constexpr (NC > 0) {  // pointer loads: 8% faster than seven buffer resources per plane (kernel 5)
        const long long pk = (long long)k * g.sz;
#pragma unroll
        for (int m = 0; m < RY; ++m)
#pragma unroll
          for (int n = 0; n < NC; ++n)
            cf[m][n] = ld_coef<T, C, STRM>(a.coef + n * a.cstride + pk + off[m] / (unsigned)sizeof(T));
      }
      // plane k+R+1 into the slot freed by plane k-R, halo ring of plane k+1
      T nq[RY];
      {
        const auto vp = pl(a.v, min(k + 1 + R, zlast));
#pragma unroll
        for (int m = 0; m < RY; ++m) nq[m] = ld(vp, off[m]);
        const auto hp = pl(a.v, min(RING ? k + 1 + R : k + 1, zlast));
#pragma unroll
        for (int s = 0; s < HS; ++s) hv[s] = ld(hp, hoff[s]);
      }

      __syncthreads();

      T *__restrict__ uk = a.u + (long long)k * g.sz;
      const bool src_plane = a.src_k >= 0 && k == a.src_k;  // src_k = -1: no source
#pragma unroll
      for (int m = 0; m < RY; ++m) {
        const int ly = R + ty * RY + m, lx = R + tx;
        auto xn = [&](int d) { return L[ly][lx + d]; };
        auto yn = [&](int d) { return L[ly + d][lx]; };
        auto zn = [&](int d) { return q[m][(p + R + d + Q) % Q]; };
        T zq[Q];
#pragma unroll
        for (int d = -R; d <= R; ++d) zq[R + d] = zn(d);
        const T v0 = zq[R];
        T out;
        if constexpr (KID == 0) {
          const T roc = ROCS ? a.roc2s : r2[m];
          out = Op<T, 0, FAST>::update(v0, uo[m], roc, Op<T, 0, FAST>::template lap<ND>(c, a.csum, v0, zq, xn, yn));
        } else if constexpr (KID == 1) {  // stencils.c:73-78
          out = c[0] * v0 + c[1] * (xn(1) + xn(-1)) + c[1] * (yn(1) + yn(-1)) + c[1] * (zq[2] + zq[0]);
        } else if constexpr (KID == 2) {  // stencils.c:94-99
          out = cf[m][0] * v0 + cf[m][1] * (xn(1) + xn(-1)) + cf[m][1] * (yn(1) + yn(-1)) +
                cf[m][1] * (zq[2] + zq[0]);
        } else if constexpr (KID == 3) {  // stencils.c:121-126
          out = cf[m][0] * v0 + cf[m][1] * (xn(1) + xn(-1)) + cf[m][2] * (yn(1) + yn(-1)) +
                cf[m][3] * (zq[2] + zq[0]);
        } else if constexpr (KID == 4) {  // stencils.c:148-162
          T l = cf[m][0] * v0;
#pragma unroll
          for (int r = 1; r <= 4; ++r) {
            l = l + cf[m][3 * r - 2] * (xn(r) + xn(-r));
            l = l + cf[m][3 * r - 1] * (yn(r) + yn(-r));
            l = l + cf[m][3 * r] * (zq[4 + r] + zq[4 - r]);
          }
          out = l;
        } else {  // KID == 5, stencils.c:193-201
          out = cf[m][0] * v0 + cf[m][1] * xn(-1) + cf[m][2] * xn(1) + cf[m][3] * yn(-1) +
                cf[m][4] * yn(1) + cf[m][5] * zq[0] + cf[m][6] * zq[2];
        }
        if (src_plane && off[m] == (unsigned)a.src_off * (unsigned)sizeof(T)) out = out + *a.src;
        if (yout[m]) {
          if constexpr (BUF)
            bstore<STRM>(out, plane_rsrc(uk, pbytes), off[m]);
          else
            stb_stream<T, STRM>(uk, off[m], out);
        }
      }
#pragma unroll
      for (int m = 0; m < RY; ++m) q[m][p] = nq[m];
    }
  }
}

}  // namespace gr
