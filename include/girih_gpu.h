/*
 * girih_gpu.h -- C ABI of the MI355X (gfx950) star-stencil time stepper.
 *
 * A drop-in GPU target for girih's time-stepper registry.  girih selects a
 * stepper as TSList[p->target_ts].func(p) with `void (*func)(Parameters *p)`
 * (src/wrappers.h:28-36, src/data_structures.h:294-297); the stepper reads
 * p->U1/U2/U3/coef in the padded layout of src/utils.c:367-374 and leaves the
 * result in p->U1 and p->U2.  This library is what such a stepper calls: the
 * host arrays go to HBM once (gr_upload), time steps run on the device
 * (gr_step), and the fields come back (gr_download).  All entry points return
 * an int status (GR_OK = 0) and never exit the process; gr_last_error()
 * describes the last failure of the calling thread.
 *
 * Plain C types only; no HIP or torch types cross this boundary.
 */
#ifndef GIRIH_GPU_H_
#define GIRIH_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_ABI_VERSION 1

/* status codes */
#define GR_OK 0
#define GR_EINVAL (-1)       /* bad argument (shape, kernel id, box, flags) */
#define GR_EHIP (-2)         /* a HIP runtime call failed */
#define GR_ENOMEM (-3)       /* device allocation failed */
#define GR_EUNSUPPORTED (-4) /* valid request this build does not implement */

/* field precision (girih's real_t: float, or double with -DDP=1, data_structures.h:99-105) */
#define GR_F64 0
#define GR_F32 1

/* gr_desc.flags */
#define GR_ROC2_SCALAR 0x1  /* kernel 0: roc2 is the constant gr_desc.roc2, no U3 array */
#define GR_FAST 0x2         /* GR_F32 kernel 0 only: neighbour pairs summed per radius before
                               scaling (radii 1-2 as differences from the centre), FMA
                               accumulation and the increment form u^n + ((u^n - u^(n-1))
                               + roc2*lap) of the leapfrog update; not bit-identical to
                               girih's operand order */
#define GR_COEF_BF16 0x4    /* roc2 / coefficient arrays stored in HBM as bfloat16
                               (computed in the field precision) */

/* Problem descriptor.  Kernel ids are girih's stencil_info_list
 * (src/kernels/stencils.c:260-269):
 *   0 25-pt star, 2nd order in time, constant coef (+ roc2 field)  "iso_ref"
 *   1  7-pt star, 1st order, constant coef                        "iso_ref_2space_1time"
 *   2  7-pt star, 1st order, variable coef (2 arrays)             "..._var"
 *   3  7-pt star, 1st order, var axis-symmetric (4 arrays)        "..._var_axsym"
 *   4 25-pt star, 1st order, var axis-symmetric (13 arrays)       "iso_ref_8space_1time_var_axsym"
 *   5  7-pt star, 1st order, var no-symmetry (7 arrays)           "..._var_nosym"      */
typedef struct gr_desc {
  int kernel;
  int dtype;        /* GR_F64 / GR_F32 */
  int nx, ny, nz;   /* interior size (girih --nx --ny --nz) */
  int flags;        /* GR_ROC2_SCALAR | GR_FAST | GR_COEF_BF16 */
  int t_block;      /* time steps fused per HBM pass: 0 or 1 = single sweeps (gr_autotune may
                       then pick a depth); a depth gr_set_config rejects -> GR_EUNSUPPORTED */
  int reserved;
  double roc2;      /* constant roc2 when GR_ROC2_SCALAR */
  double coef[11];  /* constant coefficients, kernels 0 and 1 (girih coef[0..R]) */
} gr_desc;

typedef struct gr_handle gr_handle;

typedef struct gr_stats {
  double last_ms;            /* device time of the last gr_step (HIP events) */
  int64_t last_updates;      /* lattice-site updates of the last gr_step */
  int32_t last_launches;     /* kernel launches of the last gr_step */
  int32_t t_block;           /* fused time steps per pass in use */
  int32_t tile_x, tile_y;    /* output tile of the sweep kernel, or of the fused pass when t_block > 1 */
  int32_t zchunk;            /* z planes streamed per workgroup */
  int32_t parity;            /* next application writes U1 (0) or U2 (1) */
  int64_t device_bytes;      /* HBM held by the handle */
  int32_t pitch_x, rows_y, planes_z; /* device layout of one field */
  int32_t x_offset;          /* element offset of logical x = 0 within a row */
  int32_t tile_variant;      /* gr_set_config tile variant in use */
  int32_t xcd_order;         /* workgroup order (gr_set_xcd_order): 0 plain, 1 or 2 XCD-contiguous */
} gr_stats;

/* ---- Level 1: time stepper (replaces TSList[ts].func, src/wrappers.h:28-36) ---- */

/* Allocate device fields for the problem (U1, U2, roc2 or coefficient arrays). */
int gr_create(const gr_desc *desc, gr_handle **out);
/* Free everything owned by the handle. */
int gr_destroy(gr_handle *h);
/* Describe the last error of this thread (static storage). */
const char *gr_last_error(void);

/* Copy girih host arrays to HBM.  ldomain_shape = {nnx, nny, nnz} of the host arrays
 * (girih p->ldomain_shape, utils.c:372-374; nnx may include padding); x is fastest.
 * U3 is roc2 (kernel 0 without GR_ROC2_SCALAR), coef holds girih's SoA coefficient
 * arrays (coef[m*nnx*nny*nnz + idx], stencils.h:32) for kernels 2-5.  Any pointer
 * the kernel does not use may be NULL.  Element type follows desc.dtype; with
 * GR_COEF_BF16, U3 and coef are FP32 host arrays rounded to bf16 (nearest even) on
 * the device.  Resets the step parity (the next application writes U1). */
int gr_upload(gr_handle *h, const int ldomain_shape[3], const void *U1, const void *U2,
              const void *U3, const void *coef);
/* Copy U1 and U2 back into girih-layout host arrays (either may be NULL). */
int gr_download(gr_handle *h, const int ldomain_shape[3], void *U1, void *U2);

/* Apply `napply` single time-step applications, alternating U1 <- f(U2) then
 * U2 <- f(U1) as girih's steppers do (nb_naive_ts.c:187-203).  girih's spatial
 * blocking applies 2*ceil(nt/2) applications; its diamond stepper applies
 * nt-1 after rounding nt (diamond_utils.c:1042-1056).  Synchronous. */
int gr_step(gr_handle *h, int napply);
/* girih time-stepper semantics in one call: ts 0/1 -> 2*ceil(nt/2) applications;
 * ts 2 -> t_dim/nt validated and rounded like the diamond stepper, then nt-1
 * applications.  *nt_applied (may be NULL) receives the nt girih would report. */
int gr_run_ts(gr_handle *h, int target_ts, int nt, int t_dim, int *nt_applied);

/* Point source (nb_naive_ts.c:190,196): after application s the value
 * values[s] is added at halo-inclusive point (i, j, k) of the field just written.
 * n values are copied to HBM; applications beyond n add nothing.  n = 0 clears. */
int gr_set_source(gr_handle *h, int i, int j, int k, const double *values, int n);

int gr_get_stats(gr_handle *h, gr_stats *out);

/* nt as girih's diamond stepper rounds it: (nt-2) % (2*(t_dim+1)) == 0. */
int gr_mwd_round_nt(int nt, int t_dim);

/* ---- Level 2: one sweep on device-resident girih-layout arrays ----
 * KERNEL_SIG of src/data_structures.h:202-204 with device pointers:
 * u <- f(v) over the box [xb,xe) x [yb,ye) x [zb,ze) of arrays shaped
 * shape = {nnx, nny, nnz} (x fastest).  coef: constant list (kernels 0,1, host
 * pointer to R+1 values) or SoA device arrays (kernels 2-5).  roc2: device array
 * (kernel 0).  The box must keep R cells from every array face.  Runs on the
 * default stream and synchronizes. */
int gr_sweep_dev(int kernel, int dtype, const int shape[3], int xb, int yb, int zb, int xe,
                 int ye, int ze, const void *coef, void *u, const void *v, const void *roc2);

/* Device pointer of a field for zero-copy interop: which = 0 U1, 1 U2, 2 roc2 / U3,
 * 3 coefficient arrays, 4 spare field A, 5 spare field B.  Layout via gr_get_stats:
 * element (i,j,k) at base + x_offset + (k*rows_y + j)*pitch_x + i. */
int gr_field_ptr(gr_handle *h, int which, void **dev_ptr);

/* Device-side initial states (no host traffic for large grids).
 * mode 0: girih's closed-form state (utils.c:605-697) on U1, U2 and roc2;
 * mode 1: interior of U1=U2 uniform [0,1) from `seed`, halos zero;
 * mode 2: zero fields. */
int gr_init_fields(gr_handle *h, int mode, uint64_t seed);
/* Variable coefficients from `seed`: array 0 uniform [lo0,hi0), the others
 * uniform [lo1,hi1) (kernels 2-5; whole domain incl. halo). */
int gr_init_coef_uniform(gr_handle *h, uint64_t seed, double lo0, double hi0, double lo1,
                         double hi1);

/* Launch configuration: tile variant of the sweep kernel (0 = default; see
 * gr_get_stats for the resulting tile), z planes streamed per workgroup (0 = all),
 * fused time steps per HBM pass (1 = single sweep; kernel 0: 2; kernel 1: 2..6;
 * kernel 5: 2, 4, 6; other depths return GR_EUNSUPPORTED).  Fused passes run only
 * while U1 and U2 hold the same halo (with GR_COEF_BF16: kernels 0 and 5);
 * otherwise the stepper uses single sweeps.  Values < 0 leave a setting unchanged.  Results
 * never depend on the configuration. */
int gr_set_config(gr_handle *h, int tile_variant, int zchunk, int t_block);
/* Workgroup order of the sweep kernels and fused passes: 0 = plain (the XCDs take tiles
 * round-robin); 1 = each XCD owns a contiguous range of all tiles and z chunks; 2 = each
 * XCD owns a contiguous band of the tiles of every z chunk, while all XCDs work on the
 * same chunk.  Results never depend on the order. */
int gr_set_xcd_order(gr_handle *h, int xcd_order);
/* Cache policy of the sweeps and fused passes (no counterpart in girih): mode -1
 * (default) uses non-temporal loads and stores for the data a step touches once when
 * the problem's working set exceeds 512 MB (twice the Infinity Cache), and normal
 * accesses otherwise; 0 never uses them; 1 always does.  The arithmetic is the same
 * in every mode: results never depend on it. */
int gr_set_streaming(gr_handle *h, int mode);
/* Graph mode (no counterpart in girih): gr_step captures its launch sequence into a
 * HIP graph once and replays it while the step count, parity, configuration and
 * source window repeat (one launch per call instead of one per sweep; for small,
 * launch-bound grids).  Results are identical to eager launches. */
int gr_set_graph(gr_handle *h, int on);
/* Autotuner (replaces girih's CPU tile/thread-group search, diamond_utils.c:691-847):
 * times every tile variant x workgroup order x z-chunk x fused depth candidate for
 * `steps` steps on scratch copies of the fields, times the four fastest single-sweep
 * candidates again over 4x as many steps, keeps the fastest, and reports it in *best
 * (may be NULL).  The problem's own fields and step parity are left untouched. */
int gr_autotune(gr_handle *h, int steps, gr_stats *best);

/* Global reductions in FP64 over the interior cells: field fa of handle a against
 * field fb of handle b (fields: 0 U1, 1 U2; same nx, ny, nz and kernel radius; the
 * precisions may differ).  out = { sum (a-b)^2, sum b^2, max |a-b|, max |b| }.
 * b may equal a (then fa != fb compares the two fields of one handle). */
int gr_compare_fields(gr_handle *a, int fa, gr_handle *b, int fb, double out[4]);
/* gr_compare_fields without device allocation: no counterpart in girih; for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  Same arguments, checks and results, bit for bit.  The
 * 2048 x 4 FP64 partial sums and the 4 results live in a static device array of
 * 8196 doubles (64 KiB) that the library holds on every GPU it runs on until the
 * process ends; the partial count does not depend on the grid size.  All calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int gr_compare_fields_static(gr_handle *a, int fa, gr_handle *b, int fb, double out[4]);

/* ---- Velocity models for kernel 0 (roc2 = (v*dt/h)^2, girih's U3 array) ----
 * No counterpart in girih, which fills U3 with its closed-form ramp
 * (utils.c:605-697).  A layered earth: nlayers layers along z, speeds rising
 * linearly from vmin (top) to vmax (bottom), interface depths drawn from the
 * seed; speeds scaled by 1 + pert * clamp(sum of nblobs random signed Gaussian
 * blobs, -1, 1) and by 1 + anomaly * exp(-|x - c|^2 / (2 sigma^2)).  The model is
 * a pure function of the struct and the grid size, whatever the field type. */
typedef struct gr_vmodel {
  uint64_t seed;
  int32_t nlayers;            /* 1..16 */
  int32_t nblobs;             /* 0..16 */
  double vmin, vmax;          /* m/s */
  double pert;                /* relative bound of the blob perturbation, [0, 1) */
  double dt_over_h;           /* time step / grid spacing (s/m) */
  double anomaly;             /* relative amplitude of the centred anomaly (0 = none) */
  double anomaly_center[3];   /* interior-relative cells (i-R, j-R, k-R) */
  double anomaly_sigma;       /* cells */
  int32_t round_bf16;         /* round roc2 to bfloat16 values in any storage */
  int32_t reserved;
} gr_vmodel;
int gr_init_roc2_model(gr_handle *h, const gr_vmodel *model);
/* Copy roc2 to a girih-layout host array of the field type (bf16 storage widened). */
int gr_download_roc2(gr_handle *h, const int ldomain_shape[3], void *U3);

/* ---- Batched problems (no counterpart in girih, which steps one grid) ----
 * n independent kernel-0 problems (GR_F32 | GR_FAST, roc2 field, optionally
 * GR_COEF_BF16; same flags for all) of any sizes advance together: each time step
 * (or two-step block, gr_batch_set_time_block) is ONE persistent launch whose
 * workgroups claim (problem, xy tile, z chunk) items from a device work queue,
 * largest problem first, or a whole call is one cooperative launch with a grid
 * barrier between steps (gr_batch_step mode 2).  Every problem keeps
 * its own handle (gr_batch_problem) for uploads, models, sources, downloads and
 * comparisons; its results equal gr_step's on the same handle. */
typedef struct gr_batch gr_batch;
typedef struct gr_batch_stats {
  double last_ms;           /* device time of the last gr_batch_step / tuning run */
  int64_t last_updates;     /* lattice-site updates of that call (all problems) */
  int32_t last_launches;    /* step launches of that call */
  int32_t nproblems;
  int32_t items;            /* work items per step */
  int32_t grid;             /* persistent workgroups per launch */
  int32_t variant, zchunk;  /* tile variant and z planes per item (0 = whole grid) */
  int32_t tile_x, tile_y;
  int32_t nvariants;        /* tile variants available */
  int32_t time_block;       /* configured steps per launch (1 or 2); items, grid, zchunk and
                               tile_x/y describe the two-step blocks when the last call ran them */
  int32_t block_variant;    /* two-step kernel (gr_batch_set_time_block) */
  int32_t last_mode;        /* mode of the last gr_batch_step / tuning run (0, 1 or 2) */
} gr_batch_stats;
int gr_batch_create(const gr_desc *descs, int n, gr_batch **out);
int gr_batch_destroy(gr_batch *b);
/* Handle of problem i (owned by the batch: do not gr_destroy it). */
int gr_batch_problem(gr_batch *b, int i, gr_handle **h);
/* Tile variant of single steps (0..nvariants-1) and z planes per work item
 * (0 = whole grid; single steps and two-step blocks); values < 0 leave a setting
 * unchanged.  Results never depend on them. */
int gr_batch_set_config(gr_batch *b, int variant, int zchunk);
/* Advance every problem by napply applications (girih's U1/U2 pairing per
 * problem).  use_graph selects how the steps are launched:
 *   0  one persistent launch per step (or two-step block);
 *   1  the same launches replayed from a HIP graph (captured on first use, kept
 *      while napply and the configuration stay the same);
 *   2  one cooperative persistent launch for the whole call: its workgroups run a
 *      step's work queue, then meet at a grid barrier whose agent-scope release
 *      and acquire make the step's writes visible to every XCD.  With two-step
 *      blocks an odd count adds one single-step launch.  GR_EUNSUPPORTED where
 *      the device has no cooperative launch.
 * Results do not depend on the mode. */
int gr_batch_step(gr_batch *b, int napply, int use_graph);
/* Receiver plane of problem i (shots, after nb_naive_ts.c:196's receiver, which girih
 * leaves disabled): after every application a the batch runs, the interior of plane
 * rec_k (halo-inclusive z index) of the field the step wrote goes to the device array
 * trace_dev, FP32 [max_steps][ny][nx], at trace[a] (a counts the problem's applications
 * since its fields were set, gr_shot_forward's layout); applications a >= max_steps
 * are not recorded.  trace_dev = NULL removes the receiver.  A batch with receivers
 * runs single steps (no two-step blocks). */
int gr_batch_set_receivers(gr_batch *b, int i, int rec_k, void *trace_dev, int max_steps);
/* Time steps per launch (no counterpart in girih, whose diamonds block time per
 * grid, src/kernels/diamond_ts.c): 1 = single steps; 2 = two-step blocks, where a
 * work item is (problem, 64 x 16 tile, z chunk) advanced two steps on chip, reading
 * (u^(n-1), u^n) and writing (u^(n+1), u^(n+2)) once (8 B per update in FP32
 * against 12).  A call with an odd step count ends with one single step.  Blocks
 * run only while every problem's U1 and U2 hold the same halo; otherwise a call
 * runs single steps.  Other tblock values return GR_EUNSUPPORTED.  block_variant
 * selects the two-step kernel: 0 keeps every value in registers (one workgroup
 * per CU), 1 caps them at 128 (two per CU, a few spills); < 0 leaves it unchanged.
 * Results never depend on either. */
int gr_batch_set_time_block(gr_batch *b, int tblock, int block_variant);
/* Time every tile variant x z chunk for `steps` steps, then two-step blocks per z
 * chunk, on scratch fields, launched as use_graph says (gr_batch_step), and keep
 * the fastest per step (*zchunk reports the chosen kind's chunk); the problems'
 * fields and parities are untouched. */
int gr_batch_autotune(gr_batch *b, int steps, int use_graph, int *variant, int *zchunk);
int gr_batch_get_stats(gr_batch *b, gr_batch_stats *out);

/* ---- Shots: receiver traces and the adjoint gradient (no counterpart in girih,
 * whose point source is the only I/O, nb_naive_ts.c:190,196) ----
 * A shot runs kernel 0 on handle h (per-cell roc2 field; GR_F64, or GR_F32 with
 * GR_FAST) from zero fields for nt applications with h's point source
 * (gr_set_source), recording the interior of plane rec_k (halo-inclusive z index)
 * after every application.  The gradient is taken with respect to the squared
 * slowness m = 1/v^2, roc2 = dt_over_h^2 / m, by the discrete adjoint of the
 * stepper with checkpoints every ckpt applications and recomputation.
 * Device arrays are dense, x fastest, in the field type:
 *   m, gradient [nz][ny][nx];  traces, cotangents [nt][ny][nx].
 * All calls are synchronous on the handle's stream. */
typedef struct gr_shot gr_shot;
int gr_shot_create(gr_handle *h, int rec_k, int nt, int ckpt, double dt_over_h, gr_shot **out);
int gr_shot_destroy(gr_shot *s);
/* roc2 = dt_over_h^2 / m on the whole field (halo cells copy the nearest interior m). */
int gr_set_roc2_from_slowness2(gr_handle *h, const void *m_dev, double dt_over_h);
/* Forward run: traces into trace_dev; keeps the checkpoints for gr_shot_adjoint. */
int gr_shot_forward(gr_shot *s, void *trace_dev);
/* Vector-Jacobian product of the last forward run: grad = sum_a res[a] . d trace[a] / dm. */
int gr_shot_adjoint(gr_shot *s, const void *res_dev, void *grad_dev);
/* Give dst the checkpoints of src's last forward run (same grid, nt, ckpt and
 * receiver plane), rounded to dst's field type: an FP64 forward run can seed an
 * FP32 adjoint (mixed precision: FP64 traces and residuals, FP32 recomputation and
 * adjoint), or an FP32/FP64 shot its own type's copy.  FP32 -> FP64 returns
 * GR_EUNSUPPORTED.  dst's roc2 must be set for the same m before gr_shot_adjoint. */
int gr_shot_take_checkpoints(gr_shot *dst, const gr_shot *src);

#ifdef __cplusplus
}
#endif
#endif /* GIRIH_GPU_H_ */
