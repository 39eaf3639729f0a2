// gr_batch.hip -- many independent kernel-0 wave problems of uneven size, one
// persistent launch per time step (or per two-step block), the step loop optionally
// replayed from a HIP graph or run as one cooperative launch.
//
// girih steps one grid at a time and balances its CPU thread groups with a
// dynamic ready queue of diamond tiles (src/kernels/diamond_ts.c:663-747,
// `avail_list` under omp critical).  Here the queue spans grids: the work items
// of one time step are the (grid, xy tile, z chunk) triples of every problem,
// ordered largest grid first; a grid of persistent workgroups (as many as the
// GPU holds at once) claims them with one atomic increment each, so large and
// small grids share the CUs without a launch per grid and without a tail of idle
// CUs behind the largest grid.  Each item streams its tile through its z chunk
// with the single-sweep body of gr_sweep.hpp (SweepTile::run).
//
// The claim counter is never reset between the steps of one call: step s claims
// values in [s*(items+grid), (s+1)*(items+grid)), because every workgroup makes
// exactly one failing claim per step (launches of two-step blocks and single steps
// each add their own items + grid).  One reset kernel (not a memset, so the
// sequence can live in a HIP graph) starts each call.
//
// Fields are FP32 with GR_FAST arithmetic; roc2 is a per-cell field stored as
// bf16 (GR_COEF_BF16) or FP32.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <climits>
#include <numeric>
#include <vector>

#include "../../include/girih_gpu.h"
#include "gr_internal.hpp"
#include "gr_sweep.hpp"
#include "gr_wave2.hpp"

namespace gr {
namespace {

// One problem of one step parity.  All words are 4-byte multiples (read with
// readfirstlane into scalar registers).
struct BatchDesc {
  float *u;          // written: u^(n+1) (holds u^(n-1) on entry)
  const float *v;    // read: u^n
  const void *roc2;  // bf16 or FP32 roc2 field (x offset applied)
  const float *src;  // source values of this call's first step (null = none)
  long long sy, sz;
  int xe, ye, ze;    // interior box ends (the box starts at R = 4)
  int tiles_x, tiles_y, nzc, zchunk;
  int item0;         // first work item of this problem
  int nsrc;          // source values available from `src`
  int src_k, src_off;
  float c[5], csum;
  int pad;
};
static_assert(sizeof(BatchDesc) % 8 == 0, "descriptor words");

// A problem's receiver plane: after each application the interior of plane rec_k of
// the field the step wrote goes to trace[a][j][i] (a = the problem's application
// count, gr_shot_forward's layout), for a < max_steps.
struct RecDesc {
  float *trace;
  long long plane;  // rec_k * plane size: offset of the plane in the field
  long long sy;
  int prob;         // index of the problem in the descriptor tables
  int nx, ny;
  int a0;           // applications before this call
  int max_steps;
  int row0;         // first row of this receiver in the rows of all receivers
};
static_assert(sizeof(RecDesc) % 8 == 0, "descriptor words");

// Row `row` of all receivers' planes after call step `step`, from the fields the step
// wrote (table d of that step's parity); lanes stride over x.
__device__ __forceinline__ void record_row(const BatchDesc *__restrict__ d, const RecDesc *__restrict__ r, int nrec,
                                           int row, int step, int lane, int nlanes) {
  int ri = 0;
  while (ri + 1 < nrec && r[ri + 1].row0 <= row) ++ri;
  const RecDesc R = r[ri];
  const int j = row - R.row0, a = R.a0 + step;
  if (j >= R.ny || a >= R.max_steps) return;
  const float *u = d[R.prob].u + R.plane + (long long)(j + 4) * R.sy + 4;
  float *t = R.trace + ((long long)a * R.ny + j) * R.nx;
  for (int i = lane; i < R.nx; i += nlanes) t[i] = u[i];
}

// The receivers of one step (launches per step): one workgroup per row.
__global__ __launch_bounds__(256) void batch_record_kernel(const BatchDesc *__restrict__ d,
                                                           const RecDesc *__restrict__ r, int nrec, int step) {
  record_row(d, r, nrec, blockIdx.x, step, threadIdx.x, blockDim.x);
}

constexpr int kMaxBatch = 256;

__global__ void batch_reset_kernel(unsigned long long *ctr) {
  if (threadIdx.x == 0 && threadIdx.y == 0) *ctr = 0ull;
}

template <typename X>
__device__ __forceinline__ X uniform_load(const X *p) {
  constexpr int W = sizeof(X) / 4;
  int w[W];
  const int *s = reinterpret_cast<const int *>(p);
#pragma unroll
  for (int i = 0; i < W; ++i) w[i] = __builtin_amdgcn_readfirstlane(s[i]);
  X x;
  __builtin_memcpy(&x, w, sizeof(X));
  return x;
}

// Claim and run the items of one step until the queue is empty.  `base` is the
// counter value at which the step's claims start.
template <typename C, int BX, int BY, int RY, bool RING>
__device__ __forceinline__ void step_items(const BatchDesc *__restrict__ d, int nprob, int nitems,
                                           unsigned long long *ctr, unsigned long long base, int step,
                                           typename SweepTile<float, C, 0, BY, RY, true, false, RING, kFastND, BX,
                                                              true>::Lds &lds,
                                           int *sclaim) {
  using ST = SweepTile<float, C, 0, BY, RY, true, false, RING, kFastND, BX, true>;  // a step streams every problem
  const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
  for (;;) {
    if (leader) {
      const int it = (int)(atomicAdd(ctr, 1ull) - base);
      int p = 0;
      if (it < nitems)
        while (p + 1 < nprob && d[p + 1].item0 <= it) ++p;
      sclaim[0] = it;
      sclaim[1] = p;
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(sclaim[0]);
    const int p = __builtin_amdgcn_readfirstlane(sclaim[1]);
    if (it >= nitems) break;
    const BatchDesc D = uniform_load(d + p);
    SweepArgs<float, C> a{};
    a.u = D.u;
    a.v = D.v;
    a.roc2 = (const C *)D.roc2;
    for (int r = 0; r < 5; ++r) a.c[r] = D.c[r];
    a.csum = D.csum;
    a.g.sy = D.sy;
    a.g.sz = D.sz;
    a.g.xb = a.g.yb = a.g.zb = 4;
    a.g.xe = D.xe;
    a.g.ye = D.ye;
    a.g.ze = D.ze;
    a.tiles_x = D.tiles_x;
    a.tiles_y = D.tiles_y;
    a.nzc = D.nzc;
    a.zchunk = D.zchunk;
    const bool src = D.src && step < D.nsrc;
    a.src = src ? D.src + step : nullptr;
    a.src_k = src ? D.src_k : -1;
    a.src_off = D.src_off;
    const int local = it - D.item0;
    const int txy = D.tiles_x * D.tiles_y;
    const int zc = local / txy, rem = local - zc * txy;
    const int ty = rem / D.tiles_x;
    ST::run(a, rem - ty * D.tiles_x, ty, zc, lds);
    __syncthreads();  // LDS and the claim slots are reused by the next item
  }
  __syncthreads();  // every lane has read the failing claim before the slots are reused
}

// One step per launch.
template <typename C, int BX, int BY, int RY, bool RING>
__global__ __launch_bounds__(64 * BX * BY) void batch_kernel(const BatchDesc *__restrict__ d, int nprob, int nitems,
                                                             unsigned long long *ctr, unsigned long long base,
                                                             int step) {
  using ST = SweepTile<float, C, 0, BY, RY, true, false, RING, kFastND, BX, true>;
  __shared__ typename ST::Lds lds;
  __shared__ int sclaim[2];
  step_items<C, BX, BY, RY, RING>(d, nprob, nitems, ctr, base, step, lds, sclaim);
}

// Every step of a call in one cooperative launch: the workgroups run a step's
// queue, then meet at a grid barrier.  The barrier's agent-scope release and
// acquire write back and invalidate each XCD's L2 once per step, the maintenance
// a launch boundary does, so the next step reads what every XCD wrote.  Step s
// reads table d + (s & 1) * nprob; its claims start at s * (nitems + gridDim.x).
//
// With receivers, step s's planes are recorded after its barrier, while step s+1
// runs: step s+1 writes the other field of each pair, and the barrier after it
// orders the recording before step s+2 overwrites the recorded field.
template <typename C, int BX, int BY, int RY, bool RING>
__global__ __launch_bounds__(64 * BX * BY) void batch_run_kernel(const BatchDesc *__restrict__ d, int nprob,
                                                                 int nitems, unsigned long long *ctr, int napply,
                                                                 const RecDesc *__restrict__ rec, int nrec,
                                                                 int nrows) {
  using ST = SweepTile<float, C, 0, BY, RY, true, false, RING, kFastND, BX, true>;
  __shared__ typename ST::Lds lds;
  __shared__ int sclaim[2];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const unsigned long long per = (unsigned long long)nitems + gridDim.x;
  const int lane = threadIdx.y * blockDim.x + threadIdx.x, nlanes = blockDim.x * blockDim.y;
  for (int s = 0; s < napply; ++s) {
    step_items<C, BX, BY, RY, RING>(d + (size_t)(s & 1) * nprob, nprob, nitems, ctr, per * s, s, lds, sclaim);
    grid.sync();
    for (int row = blockIdx.x; row < nrows; row += gridDim.x)
      record_row(d + (size_t)(s & 1) * nprob, rec, nrec, row, s, lane, nlanes);
  }
}

// Two-step blocks: one item is (grid, xy tile, z chunk) for two time steps, run on
// chip by the two-step pass's tile body (gr_wave2.hpp): it reads (u^(n-1), u^n) and
// writes (u^(n+1), u^(n+2)) into the problem's spare pair.  `pass` counts the blocks
// of the call; the block adds source values 2*pass and 2*pass + 1 of the call.
struct W2BatchDesc {
  W2Args<float, float> a;  // buffers of this block, tiles, z chunk, coefficients, source; a.roc2
                           // addresses the roc2 field, whose storage type the kernel instantiation names
  int item0;               // first work item of this problem
  int nsrc;                // source values available from a.src
};
static_assert(sizeof(W2BatchDesc) % 8 == 0, "descriptor words");

// The tile body's arguments with the roc2 field read as C (FP32 or bf16 storage).
template <typename C>
__device__ __forceinline__ W2Args<float, C> w2_args_as(const W2Args<float, float> &x) {
  W2Args<float, C> a;
  a.P = x.P;
  a.V = x.V;
  a.W = x.W;
  a.Y = x.Y;
  a.roc2 = reinterpret_cast<const C *>(x.roc2);
  for (int r = 0; r < 5; ++r) a.c[r] = x.c[r];
  a.csum = x.csum;
  a.roc2s = x.roc2s;
  a.sz = x.sz;
  a.sy = x.sy;
  a.nx = x.nx;
  a.ny = x.ny;
  a.nz = x.nz;
  a.tiles_x = x.tiles_x;
  a.tiles_y = x.tiles_y;
  a.nzc = x.nzc;
  a.zchunk = x.zchunk;
  a.xcd_order = x.xcd_order;
  a.src = x.src;
  a.src_k = x.src_k;
  a.src_off = x.src_off;
  return a;
}

using W2Tile = Wave2Tile<float, float, true, false, true, 1, 8, 2>;  // 64 x 16 output tiles, 512 lanes
constexpr int kW2Lanes = W2Tile::NT;

// Claim and run the items of one two-step block until the queue is empty.
template <typename C>
__device__ __forceinline__ void block_items(const W2BatchDesc *__restrict__ d, int nprob, int nitems,
                                            unsigned long long *ctr, unsigned long long base, int pass,
                                            typename Wave2Tile<float, C, true, false, true, 1, 8, 2>::Lds &lds,
                                            int *sclaim) {
  using WT = Wave2Tile<float, C, true, false, true, 1, 8, 2>;  // a block streams every problem
  const bool leader = threadIdx.x == 0;
  for (;;) {
    if (leader) {
      const int it = (int)(atomicAdd(ctr, 1ull) - base);
      int p = 0;
      if (it < nitems)
        while (p + 1 < nprob && d[p + 1].item0 <= it) ++p;
      sclaim[0] = it;
      sclaim[1] = p;
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(sclaim[0]);
    const int p = __builtin_amdgcn_readfirstlane(sclaim[1]);
    if (it >= nitems) break;
    const W2BatchDesc D = uniform_load(d + p);
    W2Args<float, C> a = w2_args_as<C>(D.a);
    const bool src = a.src && 2 * pass < D.nsrc;  // values past the end are zero padding
    a.src = src ? a.src + 2 * pass : nullptr;
    a.src_k = src ? a.src_k : -1;
    const int local = it - D.item0;
    const int ntile = a.tiles_x * a.tiles_y;
    WT::run(a, local % ntile, local / ntile, lds);
    __syncthreads();  // LDS and the claim slots are reused by the next item
  }
  __syncthreads();  // every lane has read the failing claim before the slots are reused
}

// MINW = 4 caps the registers at 128 so that two workgroups share a CU (a few
// queue values spill); MINW = 1 keeps every value in registers, one workgroup per CU.
template <typename C, int MINW>
__global__ __launch_bounds__(kW2Lanes) __attribute__((amdgpu_waves_per_eu(MINW))) void batch_w2_kernel(
    const W2BatchDesc *__restrict__ d, int nprob, int nitems, unsigned long long *ctr, unsigned long long base,
    int pass) {
  __shared__ typename Wave2Tile<float, C, true, false, true, 1, 8, 2>::Lds lds;
  __shared__ int sclaim[2];
  block_items<C>(d, nprob, nitems, ctr, base, pass, lds, sclaim);
}

// Every two-step block of a call in one cooperative launch (grid barrier between
// blocks, as batch_run_kernel).  Block p reads table 0 (p = 0), 1 (odd p) or 2
// (even p > 0) of d, or table 3 for every block when tuning.
template <typename C, int MINW>
__global__ __launch_bounds__(kW2Lanes) __attribute__((amdgpu_waves_per_eu(MINW))) void batch_w2_run_kernel(
    const W2BatchDesc *__restrict__ d, int nprob, int nitems, unsigned long long *ctr, int npasses, int tuning) {
  __shared__ typename Wave2Tile<float, C, true, false, true, 1, 8, 2>::Lds lds;
  __shared__ int sclaim[2];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const unsigned long long per = (unsigned long long)nitems + gridDim.x;
  for (int p = 0; p < npasses; ++p) {
    const int t = tuning ? 3 : p == 0 ? 0 : (p & 1) ? 1 : 2;
    block_items<C>(d + (size_t)t * nprob, nprob, nitems, ctr, per * p, p, lds, sclaim);
    grid.sync();
  }
}

typedef void (*batch_fn)(const BatchDesc *, int, int, unsigned long long *, unsigned long long, int);
typedef void (*batch_run_fn)(const BatchDesc *, int, int, unsigned long long *, int, const RecDesc *, int, int);
typedef void (*w2_fn)(const W2BatchDesc *, int, int, unsigned long long *, unsigned long long, int);
typedef void (*w2_run_fn)(const W2BatchDesc *, int, int, unsigned long long *, int, int);

// two-step kernels by register cap and roc2 storage type
constexpr int kW2Minw[] = {1, 4};
constexpr int kNumW2 = 2;
w2_fn w2_kernel(bool bf16, int v) {
  if (bf16) return v ? &batch_w2_kernel<__hip_bfloat16, 4> : &batch_w2_kernel<__hip_bfloat16, 1>;
  return v ? &batch_w2_kernel<float, 4> : &batch_w2_kernel<float, 1>;
}
w2_run_fn w2_run_kernel(bool bf16, int v) {
  if (bf16) return v ? &batch_w2_run_kernel<__hip_bfloat16, 4> : &batch_w2_run_kernel<__hip_bfloat16, 1>;
  return v ? &batch_w2_run_kernel<float, 4> : &batch_w2_run_kernel<float, 1>;
}

struct BatchVariant {
  int bx, by, ry;
  bool ring;
  batch_fn f32, bf16;          // one step per launch
  batch_run_fn run32, runbf16;  // every step of a call in one cooperative launch
};

// tiles of 64*BX x BY*RY cells, 64*BX x BY lanes
#define GR_BV(BX, BY, RY, RING)                                                                                   \
  {                                                                                                              \
    BX, BY, RY, RING, &batch_kernel<float, BX, BY, RY, RING>, &batch_kernel<__hip_bfloat16, BX, BY, RY, RING>,   \
        &batch_run_kernel<float, BX, BY, RY, RING>, &batch_run_kernel<__hip_bfloat16, BX, BY, RY, RING>         \
  }
const BatchVariant kVariants[] = {GR_BV(1, 8, 4, false), GR_BV(1, 8, 4, true), GR_BV(1, 4, 4, false),
                                  GR_BV(1, 8, 2, true),  GR_BV(1, 4, 4, true), GR_BV(2, 8, 2, true),
                                  GR_BV(2, 4, 4, true)};
#undef GR_BV
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace
}  // namespace gr

using namespace gr;

struct gr_batch {
  std::vector<gr_handle *> probs;  // in the caller's order
  std::vector<int> order;          // work order: largest grid first
  int variant = 0, zchunk = 0;
  int tblock = 1;                  // time steps per launch: 1 (single steps) or 2 (two-step blocks)
  int nitems = 0, grid = 0;
  int occupancy[kNumVariants] = {};
  int ncu = 0;
  BatchDesc *ddesc = nullptr;  // [4][n]: step parity 0/1 of the problem fields, then of the tuning spares
  std::vector<BatchDesc> hdesc;
  // two-step blocks: tables [0] first block of a call, [1] odd blocks, [2] even blocks
  // after the first, [3] tuning (reads the fields, writes the spares); [4] the single
  // step that ends an odd step count
  W2BatchDesc *dw2 = nullptr;
  std::vector<W2BatchDesc> hw2;
  int w2_zchunk = 0;   // z planes per two-step item (0 = whole grid)
  int w2_variant = 0;  // index into kW2Minw
  int w2_items = 0, w2_grid = 0, w2_occupancy[kNumW2] = {};
  int last_tblock = 1;  // steps per launch of the last call's blocks
  // one cooperative launch per call (gr_batch_step mode 2)
  bool coop = false;                                          // the device supports cooperative launches
  int run_occupancy[kNumVariants] = {}, w2_run_occupancy[kNumW2] = {};
  int run_grid = 0, w2_run_grid = 0;
  int last_mode = 0, last_grid = 0;
  // receiver planes (gr_batch_set_receivers), per problem in the caller's order
  std::vector<int> rec_k, rec_max;
  std::vector<float *> rec_trace;
  RecDesc *drec = nullptr;
  std::vector<RecDesc> hrec;
  int nrows = 0;
  unsigned long long *ctr = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool bf16 = false;
  // cached step-loop graph
  hipGraphExec_t gexec = nullptr;
  int g_napply = -1, g_variant = -1, g_zchunk = -1, g_table = -1, g_tblock = -1, g_w2_zchunk = -1, g_w2_variant = -1;
  double last_ms = 0;
  int last_launches = 0;
  long long last_updates = 0;
};

namespace {

int items_of(const Problem &P, int zchunk, int tx_cells, int ty_cells, int *tx, int *ty, int *nzc, int *zc) {
  *tx = (P.nx + tx_cells - 1) / tx_cells;
  *ty = (P.ny + ty_cells - 1) / ty_cells;
  *zc = (zchunk <= 0 || zchunk >= P.nz) ? P.nz : zchunk;
  *nzc = (P.nz + *zc - 1) / *zc;
  return *tx * *ty * *nzc;
}

// Fill the four descriptor tables: [0] / [1] advance the problems' fields from
// their parity (table 1 for odd steps of a call), [2] / [3] ping-pong between the
// two spare fields (tuning runs; no source).  The fields are those the problems
// hold after `npasses` two-step blocks (0: their current fields).
int build_tables(gr_batch *b, bool spares, int npasses = 0) {
  const int n = (int)b->probs.size();
  const BatchVariant &V = kVariants[b->variant];
  b->hdesc.assign(4 * (size_t)n, BatchDesc{});
  int item = 0;
  for (int oi = 0; oi < n; ++oi) {
    const int i = b->order[oi];
    gr_handle *h = b->probs[i];
    const Problem &P = gr_internal_problem(h);
    BatchDesc D{};
    D.sy = P.pitch;
    D.sz = (long long)P.pitch * P.rows;
    D.xe = P.nx + 4;
    D.ye = P.ny + 4;
    D.ze = P.nz + 4;
    const int cnt = items_of(P, b->zchunk, 64 * V.bx, V.by * V.ry, &D.tiles_x, &D.tiles_y, &D.nzc, &D.zchunk);
    D.item0 = item;
    item += cnt;
    D.roc2 = P.roc2;
    for (int r = 0; r < 5; ++r) D.c[r] = (float)P.c[r];
    D.csum = fast_csum(P.c);
    const int par = gr_internal_parity(h);
    int nsrc = 0;
    long long sidx = -1;
    const float *src = (const float *)gr_internal_source(h, &nsrc, &sidx);
    const long long done = gr_internal_napplied(h);
    const bool has_src = src && sidx >= 0 && done < nsrc;
    void *buf[4];  // written / read field at the problem's parity, then its spares
    if (npasses > 0)
      gr_internal_fused_buffers(h, npasses, buf);
    else {
      buf[0] = gr_internal_field(h, par);
      buf[1] = gr_internal_field(h, 1 - par);
    }
    for (int q = 0; q < 2; ++q) {
      BatchDesc E = D;
      E.u = (float *)buf[q];  // step parity q of the call writes the field read at parity 1 - q
      E.v = (const float *)buf[1 - q];
      E.src = has_src ? src + done : nullptr;
      E.nsrc = has_src ? (int)(nsrc - done) : 0;
      E.src_k = has_src ? (int)(sidx / D.sz) : -1;
      E.src_off = has_src ? (int)(sidx % D.sz) : 0;
      b->hdesc[(size_t)q * n + oi] = E;
      if (spares) {
        BatchDesc S = D;
        S.u = (float *)gr_internal_spare(h, q);
        S.v = (const float *)gr_internal_spare(h, 1 - q);
        if (!S.u || !S.v) return set_error(GR_ENOMEM, "gr_batch: cannot allocate tuning spares");
        S.src = nullptr;
        S.src_k = -1;
        b->hdesc[(size_t)(2 + q) * n + oi] = S;
      }
    }
  }
  b->hrec.clear();
  b->nrows = 0;
  for (int oi = 0; oi < n && !spares; ++oi) {  // receivers of the problems (tuning records nothing)
    const int i = b->order[oi];
    if (!b->rec_trace[i]) continue;
    const Problem &P = gr_internal_problem(b->probs[i]);
    RecDesc R{};
    R.trace = b->rec_trace[i];
    R.plane = (long long)b->rec_k[i] * P.pitch * P.rows;
    R.sy = P.pitch;
    R.prob = oi;
    R.nx = P.nx;
    R.ny = P.ny;
    R.a0 = (int)std::min<long long>(gr_internal_napplied(b->probs[i]), INT32_MAX);
    R.max_steps = b->rec_max[i];
    R.row0 = b->nrows;
    b->nrows += P.ny;
    b->hrec.push_back(R);
  }
  if (!b->hrec.empty()) {
    hipError_t e = hipMemcpyAsync(b->drec, b->hrec.data(), sizeof(RecDesc) * b->hrec.size(), hipMemcpyHostToDevice,
                                  b->stream);
    if (e != hipSuccess) return set_errorf(GR_EHIP, "gr_batch: receiver copy: %s", hipGetErrorString(e));
  }
  b->nitems = item;
  b->grid = std::max(1, std::min(item, b->ncu * std::max(1, b->occupancy[b->variant])));
  b->run_grid = std::max(1, std::min(item, b->ncu * std::max(1, b->run_occupancy[b->variant])));
  const size_t bytes = sizeof(BatchDesc) * (spares ? 4 : 2) * (size_t)n;
  hipError_t e = hipMemcpyAsync(b->ddesc, b->hdesc.data(), bytes, hipMemcpyHostToDevice, b->stream);
  if (e != hipSuccess) return set_errorf(GR_EHIP, "gr_batch: descriptor copy: %s", hipGetErrorString(e));
  return GR_OK;
}

// Fill the two-step tables: [0] the first block of a call, [1] its odd blocks,
// [2] its even blocks after the first (the buffer rotation of the single-problem
// path has period two from the second block on), [3] tuning blocks, which read the
// problems' fields and write their spares (no source).
int build_w2_tables(gr_batch *b) {
  const int n = (int)b->probs.size();
  b->hw2.assign(4 * (size_t)n, W2BatchDesc{});
  int item = 0;
  for (int oi = 0; oi < n; ++oi) {
    gr_handle *h = b->probs[b->order[oi]];
    const Problem &P = gr_internal_problem(h);
    W2BatchDesc D{};
    W2Args<float, float> &a = D.a;
    a.roc2 = (const float *)P.roc2;  // the field's address; the kernel reads it as its storage type
    for (int r = 0; r < 5; ++r) a.c[r] = (float)P.c[r];
    a.csum = fast_csum(P.c);
    a.sz = (long long)P.pitch * P.rows;
    a.sy = P.pitch;
    a.nx = P.nx;
    a.ny = P.ny;
    a.nz = P.nz;
    const int cnt = items_of(P, b->w2_zchunk, W2Tile::TX, W2Tile::TY, &a.tiles_x, &a.tiles_y, &a.nzc, &a.zchunk);
    D.item0 = item;
    item += cnt;
    int nsrc = 0;
    long long sidx = -1;
    const float *src = (const float *)gr_internal_source(h, &nsrc, &sidx);
    const long long done = gr_internal_napplied(h);
    const bool has_src = src && sidx >= 0 && done < nsrc;
    for (int t = 0; t < 4; ++t) {
      W2BatchDesc E = D;
      void *buf[4];
      gr_internal_fused_buffers(h, t < 3 ? t : 0, buf);
      E.a.P = (const float *)buf[0];  // u^(n-1)
      E.a.V = (const float *)buf[1];  // u^n
      E.a.W = (float *)buf[2];
      E.a.Y = (float *)buf[3];
      if (t == 3) {  // tuning: the fields are read, the spares written
        void *sp0 = gr_internal_spare(h, 0), *sp1 = gr_internal_spare(h, 1);
        if (!sp0 || !sp1) return set_error(GR_ENOMEM, "gr_batch: cannot allocate tuning spares");
        E.a.P = (const float *)gr_internal_field(h, 0);
        E.a.V = (const float *)gr_internal_field(h, 1);
        E.a.W = (float *)sp0;
        E.a.Y = (float *)sp1;
      }
      const bool s = has_src && t < 3;
      E.a.src = s ? src + done : nullptr;
      E.a.src_k = s ? (int)(sidx / a.sz) : -1;
      E.a.src_off = s ? (int)(sidx % a.sz) : 0;
      E.nsrc = s ? (int)(nsrc - done) : 0;
      b->hw2[(size_t)t * n + oi] = E;
    }
  }
  b->w2_items = item;
  b->w2_grid = std::max(1, std::min(item, b->ncu * std::max(1, b->w2_occupancy[b->w2_variant])));
  b->w2_run_grid = std::max(1, std::min(item, b->ncu * std::max(1, b->w2_run_occupancy[b->w2_variant])));
  hipError_t e = hipMemcpyAsync(b->dw2, b->hw2.data(), sizeof(W2BatchDesc) * b->hw2.size(),
                                hipMemcpyHostToDevice, b->stream);
  if (e != hipSuccess) return set_errorf(GR_EHIP, "gr_batch: descriptor copy: %s", hipGetErrorString(e));
  return GR_OK;
}

// Enqueue one call: napply single steps from table `table` (0: the problems, 2:
// spares), or with tblock == 2, napply / 2 two-step blocks (table 3 of the
// two-step tables when tuning) and, for an odd count, one single step from table 0
// (built for the fields after the blocks).  The claim counter's base advances by
// items + grid per launch.
hipError_t enqueue_steps(gr_batch *b, int table, int napply, int tblock, bool tuning, hipStream_t st) {
  const int n = (int)b->probs.size();
  const BatchVariant &V = kVariants[b->variant];
  const batch_fn f = b->bf16 ? V.bf16 : V.f32;
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, b->ctr);
  unsigned long long base = 0;
  int s = 0;
  if (tblock == 2) {
    for (int pass = 0; 2 * pass + 1 < napply; ++pass, s += 2) {
      const int t = tuning ? 3 : pass == 0 ? 0 : (pass & 1) ? 1 : 2;
      const W2BatchDesc *tab = b->dw2 + (size_t)t * n;
      hipLaunchKernelGGL(w2_kernel(b->bf16, b->w2_variant), dim3(b->w2_grid), dim3(kW2Lanes), 0, st, tab, n,
                         b->w2_items, b->ctr, base, pass);
      base += (unsigned long long)(b->w2_items + b->w2_grid);
    }
    table = 0;
  }
  const int nrec = table == 0 ? (int)b->hrec.size() : 0;
  for (int q = 0; s < napply; ++s, ++q) {
    const BatchDesc *tab = b->ddesc + (size_t)(table + (q & 1)) * n;
    hipLaunchKernelGGL(f, dim3(b->grid), dim3(64 * V.bx, V.by), 0, st, tab, n, b->nitems, b->ctr, base, s);
    base += (unsigned long long)(b->nitems + b->grid);
    if (nrec) hipLaunchKernelGGL(batch_record_kernel, dim3(b->nrows), dim3(256), 0, st, tab, b->drec, nrec, s);
  }
  return hipGetLastError();
}

// enqueue_steps' work as one cooperative launch: every single step of the call, or
// every two-step block (then, for an odd count, one single-step launch whose claims
// start after the blocks').  Never captured into a graph.
hipError_t enqueue_run(gr_batch *b, int table, int napply, int tblock, bool tuning, hipStream_t st) {
  int n = (int)b->probs.size();
  const BatchVariant &V = kVariants[b->variant];
  unsigned long long *ctr = b->ctr;
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, b->ctr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (tblock != 2) {
    const BatchDesc *tab = b->ddesc + (size_t)table * n;
    int items = b->nitems;
    const RecDesc *rec = b->drec;
    int nrec = table == 0 ? (int)b->hrec.size() : 0, nrows = nrec ? b->nrows : 0;
    void *args[] = {&tab, &n, &items, &ctr, &napply, &rec, &nrec, &nrows};
    const void *f = reinterpret_cast<const void *>(b->bf16 ? V.runbf16 : V.run32);
    return hipLaunchCooperativeKernel(f, dim3(b->run_grid), dim3(64 * V.bx, V.by), args, 0, st);
  }
  int npasses = napply / 2, tun = tuning ? 1 : 0, items = b->w2_items;
  const W2BatchDesc *tab = b->dw2;
  void *args[] = {&tab, &n, &items, &ctr, &npasses, &tun};
  const void *f = reinterpret_cast<const void *>(w2_run_kernel(b->bf16, b->w2_variant));
  e = hipLaunchCooperativeKernel(f, dim3(b->w2_run_grid), dim3(kW2Lanes), args, 0, st);
  if (e != hipSuccess || !(napply & 1)) return e;
  const unsigned long long base = (unsigned long long)npasses * (b->w2_items + b->w2_run_grid);
  hipLaunchKernelGGL(b->bf16 ? V.bf16 : V.f32, dim3(b->grid), dim3(64 * V.bx, V.by), 0, st, b->ddesc, n, b->nitems,
                     b->ctr, base, 2 * npasses);
  return hipGetLastError();
}

void drop_graph(gr_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->g_napply = b->g_variant = b->g_zchunk = b->g_table = b->g_tblock = b->g_w2_zchunk = b->g_w2_variant = -1;
}

// Run one call (enqueue_steps' arguments), timed: mode 0 launches per step (or
// block), 1 replays them from a graph, 2 runs them as one cooperative launch.
int run_steps(gr_batch *b, int table, int napply, int mode, int tblock = 1, bool tuning = false) {
  hipError_t e = hipSuccess;
  const int gtable = tuning ? 3 : table;
  const bool graph = mode == 1;
  if (mode == 2 && napply > 0) {
    e = hipEventRecord(b->ev0, b->stream);
    if (e == hipSuccess) e = enqueue_run(b, table, napply, tblock, tuning, b->stream);
  } else if (graph && napply > 0) {
    if (!b->gexec || b->g_napply != napply || b->g_variant != b->variant || b->g_zchunk != b->zchunk ||
        b->g_table != gtable || b->g_tblock != tblock || b->g_w2_zchunk != b->w2_zchunk ||
        b->g_w2_variant != b->w2_variant) {
      drop_graph(b);
      hipGraph_t g = nullptr;
      e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t e2 = enqueue_steps(b, table, napply, tblock, tuning, b->stream);
        e = hipStreamEndCapture(b->stream, &g);
        if (e == hipSuccess) e = e2;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        drop_graph(b);
        return set_errorf(GR_EHIP, "gr_batch: graph capture: %s", hipGetErrorString(e));
      }
      b->g_napply = napply;
      b->g_variant = b->variant;
      b->g_zchunk = b->zchunk;
      b->g_table = gtable;
      b->g_tblock = tblock;
      b->g_w2_zchunk = b->w2_zchunk;
      b->g_w2_variant = b->w2_variant;
    }
    e = hipEventRecord(b->ev0, b->stream);
    if (e == hipSuccess) e = hipGraphLaunch(b->gexec, b->stream);
  } else {
    e = hipEventRecord(b->ev0, b->stream);
    if (e == hipSuccess) e = enqueue_steps(b, table, napply, tblock, tuning, b->stream);
  }
  if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
  if (e != hipSuccess) return set_errorf(GR_EHIP, "gr_batch: step launch: %s", hipGetErrorString(e));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  b->last_ms = ms;
  b->last_tblock = tblock;
  b->last_mode = mode;
  if (mode == 2) {
    b->last_launches = tblock == 2 ? (napply >= 2) + (napply & 1) : (napply > 0);
    b->last_grid = tblock == 2 ? b->w2_run_grid : b->run_grid;
  } else {
    b->last_launches = tblock == 2 ? napply / 2 + (napply & 1) : napply;
    b->last_grid = tblock == 2 ? b->w2_grid : b->grid;
  }
  long long up = 0;
  for (gr_handle *h : b->probs) {
    const Problem &P = gr_internal_problem(h);
    up += (long long)P.nx * P.ny * P.nz;
  }
  b->last_updates = up * napply;
  return GR_OK;
}

// Whether every problem may run two-step blocks now (fields with equal halos;
// spares allocated and holding that halo).
int fused_ready(gr_batch *b, bool *ok) {
  *ok = b->w2_occupancy[b->w2_variant] >= 1;
  for (float *t : b->rec_trace) *ok = *ok && !t;  // receivers are recorded after single steps only
  for (gr_handle *h : b->probs) {
    if (!*ok) break;
    const int rc = gr_internal_fused_prepare(h, ok);
    if (rc) return rc;
  }
  return GR_OK;
}

int sync_problems(gr_batch *b) {
  for (gr_handle *h : b->probs) {
    const hipError_t e = hipStreamSynchronize(gr_internal_stream(h));
    if (e != hipSuccess) return set_errorf(GR_EHIP, "gr_batch: %s", hipGetErrorString(e));
  }
  return GR_OK;
}

}  // namespace

extern "C" int gr_batch_destroy(gr_batch *b) {
  if (!b) return GR_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (gr_handle *h : b->probs) gr_destroy(h);
  if (b->ddesc) (void)hipFree(b->ddesc);
  if (b->dw2) (void)hipFree(b->dw2);
  if (b->drec) (void)hipFree(b->drec);
  if (b->ctr) (void)hipFree(b->ctr);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GR_OK;
}

extern "C" int gr_batch_create(const gr_desc *descs, int n, gr_batch **out) {
  if (!descs || !out) return set_error(GR_EINVAL, "gr_batch_create: null argument");
  *out = nullptr;
  if (n < 1 || n > kMaxBatch) return set_errorf(GR_EINVAL, "gr_batch_create: need 1..%d problems", kMaxBatch);
  const int flags0 = descs[0].flags;
  for (int i = 0; i < n; ++i) {
    const gr_desc &d = descs[i];
    if (d.kernel != 0 || d.dtype != GR_F32 || !(d.flags & GR_FAST) || (d.flags & GR_ROC2_SCALAR))
      return set_errorf(GR_EUNSUPPORTED,
                        "gr_batch_create: problem %d: batches hold kernel 0, GR_F32 | GR_FAST problems with a "
                        "roc2 field", i);
    if (d.flags != flags0) return set_error(GR_EINVAL, "gr_batch_create: all problems need the same flags");
  }
  gr_batch *b = new gr_batch();
  b->bf16 = (flags0 & GR_COEF_BF16) != 0;
  b->rec_k.assign(n, 0);
  b->rec_max.assign(n, 0);
  b->rec_trace.assign(n, nullptr);
  int rc = GR_OK;
  for (int i = 0; i < n && rc == GR_OK; ++i) {
    gr_handle *h = nullptr;
    rc = gr_create(&descs[i], &h);
    if (rc == GR_OK) b->probs.push_back(h);
  }
  hipError_t e = hipSuccess;
  if (rc == GR_OK) {
    int dev = 0;
    e = hipMalloc(&b->ddesc, sizeof(BatchDesc) * 4 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&b->dw2, sizeof(W2BatchDesc) * 4 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&b->drec, sizeof(RecDesc) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&b->ctr, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int coop = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    b->coop = coop != 0;
    for (int v = 0; v < kNumVariants && e == hipSuccess; ++v) {
      const BatchVariant &V = kVariants[v];
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &b->occupancy[v], reinterpret_cast<const void *>(b->bf16 ? V.bf16 : V.f32), 64 * V.bx * V.by, 0);
    }
    for (int v = 0; v < kNumVariants && e == hipSuccess; ++v) {
      const BatchVariant &V = kVariants[v];
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &b->run_occupancy[v], reinterpret_cast<const void *>(b->bf16 ? V.runbf16 : V.run32), 64 * V.bx * V.by, 0);
    }
    for (int v = 0; v < kNumW2 && e == hipSuccess; ++v) {
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->w2_occupancy[v],
                                                       reinterpret_cast<const void *>(w2_kernel(b->bf16, v)),
                                                       kW2Lanes, 0);
      if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b->w2_run_occupancy[v], reinterpret_cast<const void *>(w2_run_kernel(b->bf16, v)), kW2Lanes, 0);
    }
    if (e == hipSuccess && b->w2_occupancy[0] < 1) b->w2_variant = 1;
    if (e != hipSuccess) rc = set_errorf(GR_EHIP, "gr_batch_create: %s", hipGetErrorString(e));
  }
  if (rc != GR_OK) {
    gr_batch_destroy(b);
    return rc;
  }
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
    const Problem &A = gr_internal_problem(b->probs[x]), &B = gr_internal_problem(b->probs[y]);
    return (long long)A.nx * A.ny * A.nz > (long long)B.nx * B.ny * B.nz;
  });
  *out = b;
  return GR_OK;
}

extern "C" int gr_batch_problem(gr_batch *b, int i, gr_handle **h) {
  if (!b || !h) return set_error(GR_EINVAL, "gr_batch_problem: null argument");
  if (i < 0 || i >= (int)b->probs.size()) return set_errorf(GR_EINVAL, "gr_batch_problem: no problem %d", i);
  *h = b->probs[i];
  return GR_OK;
}

extern "C" int gr_batch_set_config(gr_batch *b, int variant, int zchunk) {
  if (!b) return set_error(GR_EINVAL, "gr_batch_set_config: null batch");
  if (variant >= kNumVariants)
    return set_errorf(GR_EINVAL, "gr_batch_set_config: variant %d (0..%d)", variant, kNumVariants - 1);
  if (variant >= 0) b->variant = variant;
  if (zchunk >= 0) b->zchunk = b->w2_zchunk = zchunk;
  return GR_OK;
}

namespace {

// mode 2 needs cooperative launches and, for the configured kernels, workgroups that fit a CU
int check_mode(gr_batch *b, int mode, const char *who) {
  if (mode < 0 || mode > 2) return set_errorf(GR_EINVAL, "%s: mode %d (0 launches, 1 graph, 2 one launch)", who, mode);
  if (mode == 2 && !b->coop)
    return set_errorf(GR_EUNSUPPORTED, "%s: the device does not support cooperative launches", who);
  return GR_OK;
}

}  // namespace

extern "C" int gr_batch_step(gr_batch *b, int napply, int use_graph) {
  if (!b) return set_error(GR_EINVAL, "gr_batch_step: null batch");
  if (napply < 0) return set_error(GR_EINVAL, "gr_batch_step: negative step count");
  int rc = check_mode(b, use_graph, "gr_batch_step");
  if (rc) return rc;
  if (use_graph == 2 && b->run_occupancy[b->variant] < 1)
    return set_error(GR_EUNSUPPORTED, "gr_batch_step: the one-launch kernel of this tile variant does not fit a CU");
  rc = sync_problems(b);
  if (rc) return rc;
  bool fuse = false;
  if (b->tblock == 2 && napply >= 2 && (rc = fused_ready(b, &fuse))) return rc;
  const int npasses = fuse ? napply / 2 : 0;
  if (fuse) rc = build_w2_tables(b);
  if (rc == GR_OK) rc = build_tables(b, false, npasses);
  if (rc == GR_OK && fuse && use_graph == 2 && b->w2_run_occupancy[b->w2_variant] < 1)
    rc = set_error(GR_EUNSUPPORTED, "gr_batch_step: the one-launch two-step kernel does not fit a CU");
  if (rc == GR_OK) rc = run_steps(b, 0, napply, use_graph, fuse ? 2 : 1);
  if (rc != GR_OK) return rc;
  for (gr_handle *h : b->probs) {
    if (fuse)
      gr_internal_fused_advance(h, npasses, napply, b->last_ms);
    else
      gr_internal_advance(h, napply, b->last_ms);
  }
  return GR_OK;
}

extern "C" int gr_batch_set_time_block(gr_batch *b, int tblock, int block_variant) {
  if (!b) return set_error(GR_EINVAL, "gr_batch_set_time_block: null batch");
  if (tblock != 1 && tblock != 2)
    return set_errorf(GR_EUNSUPPORTED, "gr_batch_set_time_block: %d steps per launch (1 or 2)", tblock);
  if (block_variant >= kNumW2)
    return set_errorf(GR_EINVAL, "gr_batch_set_time_block: block variant %d (0..%d)", block_variant, kNumW2 - 1);
  const int wv = block_variant >= 0 ? block_variant : b->w2_variant;
  if (tblock == 2 && b->w2_occupancy[wv] < 1)
    return set_error(GR_EUNSUPPORTED, "gr_batch_set_time_block: the two-step kernel does not fit a CU");
  b->tblock = tblock;
  b->w2_variant = wv;
  return GR_OK;
}

extern "C" int gr_batch_set_receivers(gr_batch *b, int i, int rec_k, void *trace_dev, int max_steps) {
  if (!b) return set_error(GR_EINVAL, "gr_batch_set_receivers: null batch");
  if (i < 0 || i >= (int)b->probs.size()) return set_errorf(GR_EINVAL, "gr_batch_set_receivers: no problem %d", i);
  const Problem &P = gr_internal_problem(b->probs[i]);
  if (trace_dev && (rec_k < P.R || rec_k >= P.nz + P.R))
    return set_errorf(GR_EINVAL, "gr_batch_set_receivers: plane %d is not an interior plane", rec_k);
  if (trace_dev && max_steps < 0) return set_error(GR_EINVAL, "gr_batch_set_receivers: negative max_steps");
  drop_graph(b);  // the launch sequence records another set of planes
  b->rec_trace[i] = (float *)trace_dev;
  b->rec_k[i] = trace_dev ? rec_k : 0;
  b->rec_max[i] = trace_dev ? max_steps : 0;
  return GR_OK;
}

extern "C" int gr_batch_autotune(gr_batch *b, int steps, int use_graph, int *variant, int *zchunk) {
  if (!b) return set_error(GR_EINVAL, "gr_batch_autotune: null batch");
  if (steps < 1) steps = 4;
  if (int mrc = check_mode(b, use_graph, "gr_batch_autotune")) return mrc;
  const int warm = use_graph == 2 ? 2 : 0;  // the warm-up loads the timed kernel's code object
  for (gr_handle *h : b->probs)  // allocate the spares up front: their set-up runs on the problems' streams
    if (!gr_internal_spare(h, 0)) return set_error(GR_ENOMEM, "gr_batch: cannot allocate tuning spares");
  int rc = sync_problems(b);
  if (rc) return rc;
  const int zcands[] = {0, 256, 128, 64};
  double best = 1e30;  // ms per step
  int bv = b->variant, bz = b->zchunk;
  for (int v = 0; v < kNumVariants; ++v) {
    if (b->occupancy[v] < 1 || (use_graph == 2 && b->run_occupancy[v] < 1)) continue;
    for (int zc : zcands) {
      b->variant = v;
      b->zchunk = zc;
      if ((rc = build_tables(b, true))) return rc;
      if ((rc = run_steps(b, 2, 1, warm))) return rc;  // warm-up
      if ((rc = run_steps(b, 2, steps, use_graph))) return rc;
      if (b->last_ms / steps < best) {
        best = b->last_ms / steps;
        bv = v;
        bz = zc;
      }
    }
  }
  b->variant = bv;
  b->zchunk = bz;
  // two-step blocks, timed per step against the best single step
  int btb = 1, bwz = b->w2_zchunk, bwv = b->w2_variant;
  bool fuse = false;
  if ((rc = fused_ready(b, &fuse))) return rc;
  const int passes = std::max(1, steps / 2);
  for (int wv = 0; wv < kNumW2 && fuse; ++wv) {
    if (b->w2_occupancy[wv] < 1 || (use_graph == 2 && b->w2_run_occupancy[wv] < 1)) continue;
    for (int zc : zcands) {
      b->w2_variant = wv;
      b->w2_zchunk = zc;
      if ((rc = build_w2_tables(b))) return rc;
      if ((rc = run_steps(b, 0, 2, warm, 2, true))) return rc;  // warm-up
      if ((rc = run_steps(b, 0, 2 * passes, use_graph, 2, true))) return rc;
      if (b->last_ms / (2 * passes) < best) {
        best = b->last_ms / (2 * passes);
        btb = 2;
        bwz = zc;
        bwv = wv;
      }
    }
  }
  drop_graph(b);
  b->tblock = btb;
  b->w2_zchunk = bwz;
  b->w2_variant = bwv;
  if (variant) *variant = bv;
  if (zchunk) *zchunk = btb == 2 ? bwz : bz;
  return GR_OK;
}

extern "C" int gr_batch_get_stats(gr_batch *b, gr_batch_stats *out) {
  if (!b || !out) return set_error(GR_EINVAL, "gr_batch_get_stats: null argument");
  gr_batch_stats s{};
  s.last_ms = b->last_ms;
  s.last_updates = b->last_updates;
  s.last_launches = b->last_launches;
  s.nproblems = (int)b->probs.size();
  s.items = b->nitems;
  s.grid = b->last_grid > 0 ? b->last_grid : b->grid;
  s.variant = b->variant;
  s.zchunk = b->zchunk;
  s.tile_x = 64 * kVariants[b->variant].bx;
  s.tile_y = kVariants[b->variant].by * kVariants[b->variant].ry;
  s.nvariants = kNumVariants;
  s.time_block = b->tblock;
  s.block_variant = b->w2_variant;
  s.last_mode = b->last_mode;
  if (b->last_tblock == 2) {  // the last call ran two-step blocks: their items and tiles
    s.items = b->w2_items;
    s.zchunk = b->w2_zchunk;
    s.tile_x = W2Tile::TX;
    s.tile_y = W2Tile::TY;
  }
  *out = s;
  return GR_OK;
}
