/*
 * mwd_kernel_gpu -- girih's mwd_kernel command line, with the time steps run on
 * one MI355X through libgirih_gpu (include/girih_gpu.h).
 *
 * Keeps the driver flow of src/driver.c:21-82: param_default -> parse_args ->
 * init -> verify | performance_test, the flag set of src/utils.c:1227-1263, the
 * defaults of src/utils.c:41-123, and the "Key: value" report lines of
 * src/utils.c:866-1065 that girih's scripts/parse.py scrapes.
 *
 * Time steppers (TSList, src/wrappers.h:28-36):
 *   0 "Spatial Blocking", 1 "Halo-first": girih's step pairing, 2*ceil(nt/2)
 *       applications, U1 = step N-1, U2 = step N;
 *   2 "Diamond": the diamond stepper's validity rules and nt rounding
 *       (diamond_utils.c:1023-1098), nt-1 applications (U1 = step nt-1);
 *   3 "MI355X": spatial-blocking semantics with the autotuned GPU launch.
 * All four run on the GPU; 0-2 reproduce girih's CPU results bit for bit.
 * --mwd-type, --thread-group-size, --thx/--thy/--thz/--thc, --threads,
 * --cache-size, --num-wavefronts are CPU threading knobs: they are parsed and
 * checked like girih does and otherwise have no effect.
 *
 * Precision follows the build like girih's (Makefile:15-19): -DDP=1 gives the
 * FP64 binary, otherwise FP32.  --fast 1 (FP32, kernel 0) selects the
 * compensated FP32 arithmetic.
 *
 * --verify 1 runs a serial CPU evaluation of the same recurrence (girih's
 * verification.c:52-312 loop, operator order of stencils.c) and compares exactly
 * (verification.c:823-860).
 */
#define _POSIX_C_SOURCE 200809L
#include <getopt.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "girih_gpu.h"

#ifndef DP
#define DP 0
#endif
#if DP
typedef double real_t;
#define GR_DTYPE GR_F64
#else
typedef float real_t;
#define GR_DTYPE GR_F32
#endif

#define BOUNDARY_SRC_VAL (100.1)

/* stencil_info_list (src/kernels/stencils.c:260-269), regular star kernels */
static const struct {
  const char *name;
  int r, time_order, ncoef;
  const char *coeff;
} KINFO[6] = {
    {"star", 4, 2, 0, "constant"},
    {"star", 1, 1, 0, "constant"},
    {"star", 1, 1, 2, "variable"},
    {"star", 1, 1, 4, "variable axis-symmetric"},
    {"star", 4, 1, 13, "variable axis-symmetric"},
    {"star", 1, 1, 7, "variable no-symmetry"},
};
static const char *TS_NAME[] = {"Spatial Blocking", "Halo-first", "Diamond", "MI355X", 0};
static const char *MWD_NAME[] = {"Wavefront", "Fixed execution wavefronts", "Relaxed synchronization wavefront",
                                 "Relaxed synchronization wavefront with fixed execution", 0};

typedef struct {
  int nx, ny, nz, nt, n_tests, alignment, verbose, debug, verify, target_ts, target_kernel;
  int t_dim, mwd_type, thread_group_size, thx, thy, thz, thc, num_wf, cache_size, halo_concat;
  int fast, nt_applied, R;
  int ld[3]; /* padded local domain {nnx, nny, nnz} */
  uint64_t ln_domain, ln_stencils;
  real_t *U1, *U2, *U3, *coef;
  real_t g_coef[11];
} Params;

static void die(const char *msg) {
  fprintf(stderr, "ERROR: %s\n", msg);
  exit(1);
}

/* utils.c:41-123 */
static void param_default(Params *p) {
  static const double g[11] = {-0.28472, 0.16000, -0.02000, 0.00254, -0.00018, -0.18472,
                               0.19,     -0.0500, 0.00554,  -0.0009, 0.00354};
  memset(p, 0, sizeof(*p));
  p->nx = 256;
  p->ny = 64;
  p->nz = 64;
  p->alignment = 8;
  p->n_tests = 3;
  p->nt = 100;
  p->verbose = 1;
  p->t_dim = -1;
  p->halo_concat = 1;
  p->thread_group_size = p->thx = p->thy = p->thz = p->thc = -1;
  p->num_wf = -1;
  for (int i = 0; i < 11; i++) p->g_coef[i] = (real_t)g[i];
}

static void list_kernels(void) {
  printf("Available time steppers:\n#    Name\n");
  for (int i = 0; TS_NAME[i]; i++) printf("%02d   %s\n", i, TS_NAME[i]);
  printf("\nAvailable stencil kernels:\n");
  for (int i = 0; i < 6; i++)
    printf("%02d  stencil_op:%s  time-order:%d  radius:%d  coeff:%s\n", i, KINFO[i].name, KINFO[i].time_order,
           KINFO[i].r, KINFO[i].coeff);
  printf("\nAvailable MWD implementations:\n#    Name\n");
  for (int i = 0; MWD_NAME[i]; i++) printf("%02d   %s\n", i, MWD_NAME[i]);
  exit(0);
}

static void print_help(void) {
  printf("Usage (girih mwd_kernel flags; time steps run on the GPU):\n"
         "  --help | --list | --verify <bool>\n"
         "  --target-ts <0 SB | 1 Halo-first | 2 Diamond | 3 MI355X>  --target-kernel <0-5>\n"
         "  --nx --ny --nz --nt <int>  --n-tests <int>  --alignment <int>  --verbose <bool>  --debug <bool>\n"
         "  --t-dim <odd int> (diamond)  --mwd-type <0-3>  --thread-group-size --thx --thy --thz --thc\n"
         "  --num-wavefronts --cache-size --threads --wavefront --halo-concatenate --z-mpi-contig\n"
         "  --npx --npy --npz (must be 1: one GPU)  --pad-array --use-omp-stat-sched --disable-source-point\n"
         "  --fast <bool> (FP32 kernel 0: compensated arithmetic, not bit-identical)\n");
  exit(0);
}

/* utils.c:1219-1356 */
static void parse_args(int argc, char **argv, Params *p) {
  static struct option lo[] = {
      {"nz", 1, 0, 0}, {"ny", 1, 0, 0}, {"nx", 1, 0, 0}, {"nt", 1, 0, 0}, {"alignment", 1, 0, 0},
      {"verbose", 1, 0, 0}, {"debug", 1, 0, 0}, {"target-ts", 1, 0, 0}, {"target-kernel", 1, 0, 0},
      {"n-tests", 1, 0, 0}, {"verify", 1, 0, 0}, {"list", 0, 0, 0}, {"help", 0, 0, 0}, {"npx", 1, 0, 0},
      {"npy", 1, 0, 0}, {"npz", 1, 0, 0}, {"t-dim", 1, 0, 0}, {"z-mpi-contig", 1, 0, 0},
      {"disable-source-point", 0, 0, 0}, {"halo-concatenate", 1, 0, 0}, {"thread-group-size", 1, 0, 0},
      {"cache-size", 1, 0, 0}, {"wavefront", 1, 0, 0}, {"num-wavefronts", 1, 0, 0}, {"pad-array", 0, 0, 0},
      {"mwd-type", 1, 0, 0}, {"thx", 1, 0, 0}, {"thy", 1, 0, 0}, {"thz", 1, 0, 0}, {"thc", 1, 0, 0},
      {"threads", 1, 0, 0}, {"use-omp-stat-sched", 0, 0, 0}, {"fast", 1, 0, 0}, {0, 0, 0, 0}};
  int c, idx;
  while ((c = getopt_long(argc, argv, "", lo, &idx)) != -1) {
    if (c != 0) {
      fprintf(stderr, "Invalid arguments\n\n");
      print_help();
    }
    const char *n = lo[idx].name;
    int v = optarg ? atoi(optarg) : 0;
    if (!strcmp(n, "nz")) p->nz = v;
    else if (!strcmp(n, "ny")) p->ny = v;
    else if (!strcmp(n, "nx")) p->nx = v;
    else if (!strcmp(n, "nt")) p->nt = v;
    else if (!strcmp(n, "alignment")) p->alignment = v;
    else if (!strcmp(n, "verbose")) p->verbose = v != 0;
    else if (!strcmp(n, "debug")) p->debug = v != 0;
    else if (!strcmp(n, "target-ts")) p->target_ts = v;
    else if (!strcmp(n, "target-kernel")) p->target_kernel = v;
    else if (!strcmp(n, "n-tests")) p->n_tests = v;
    else if (!strcmp(n, "verify")) p->verify = v != 0;
    else if (!strcmp(n, "list")) list_kernels();
    else if (!strcmp(n, "help")) print_help();
    else if (!strcmp(n, "npx") || !strcmp(n, "npy") || !strcmp(n, "npz")) {
      if (v != 1) die("requested MPI topology shape does not match the available processes count (one GPU)");
    } else if (!strcmp(n, "t-dim")) p->t_dim = v;
    else if (!strcmp(n, "halo-concatenate")) p->halo_concat = v != 0;
    else if (!strcmp(n, "thread-group-size")) p->thread_group_size = v;
    else if (!strcmp(n, "cache-size")) p->cache_size = v;
    else if (!strcmp(n, "num-wavefronts")) p->num_wf = v;
    else if (!strcmp(n, "mwd-type")) p->mwd_type = v;
    else if (!strcmp(n, "thx")) p->thx = v;
    else if (!strcmp(n, "thy")) p->thy = v;
    else if (!strcmp(n, "thz")) p->thz = v;
    else if (!strcmp(n, "thc")) p->thc = v;
    else if (!strcmp(n, "fast")) p->fast = v != 0;
    /* z-mpi-contig, disable-source-point, wavefront, pad-array, threads,
       use-omp-stat-sched: accepted, no effect on one GPU */
  }
  if (optind < argc) {
    fprintf(stderr, "Invalid arguments\n\n");
    print_help();
  }
}

/* utils.c:317-433 (one rank) plus the diamond checks of diamond_utils.c:850-1099 */
static void init(Params *p) {
  if (p->target_kernel < 0 || p->target_kernel > 5) die("kernel not available on the GPU target (0-5)");
  if (p->target_ts < 0 || p->target_ts > 3) die("unknown time stepper");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) die("empty domain");
  p->R = KINFO[p->target_kernel].r;
  p->halo_concat = 0; /* single process (utils.c:331-334) */
  int nnx = p->nx + 2 * p->R, pad = 0;
  if (p->alignment > 0 && nnx % p->alignment) pad = p->alignment - nnx % p->alignment;
  p->ld[0] = nnx + pad;
  p->ld[1] = p->ny + 2 * p->R;
  p->ld[2] = p->nz + 2 * p->R;
  p->ln_domain = (uint64_t)p->ld[0] * p->ld[1] * p->ld[2];
  p->ln_stencils = (uint64_t)p->nx * p->ny * p->nz;
  if (p->fast && (DP || p->target_kernel != 0)) die("--fast applies to the FP32 build and kernel 0");
  if (p->target_ts == 2) {
    if ((p->mwd_type > 1) && (p->R > 1))
      die("Relaxed synchronization implementations are disabled for stencil radius > 1 for performance reasons");
    if (p->t_dim == -1) p->t_dim = 1; /* girih autotunes t_dim; the GPU does not need it */
    if (p->thread_group_size == 1) p->thx = p->thy = p->thz = p->thc = 1;
    if (p->num_wf == -1) p->num_wf = 1;
    /* diamond_utils.c:1042-1056: nt rounded (and reported) at initialisation */
    const int nt2 = gr_mwd_round_nt(p->nt, p->t_dim);
    if (nt2 != p->nt) {
      if (p->verbose) printf("###INFO: Modified nt from %03d to %03d for the intra-diamond method to work properly\n", p->nt, nt2);
      p->nt = nt2;
    }
  }
}

/* utils.c:436-496 */
static void init_coeff(Params *p) {
  const uint64_t n = p->ln_domain;
  const int R = p->R, k = p->target_kernel;
  if (KINFO[k].ncoef == 0) {
    for (int i = 0; i <= R; i++) p->coef[i] = p->g_coef[i];
    return;
  }
  if (k == 2) {
    for (int m = 0; m <= R; m++)
      for (uint64_t i = 0; i < n; i++) p->coef[i + m * n] = p->g_coef[m];
  } else if (k == 3 || k == 4) {
    for (uint64_t i = 0; i < n; i++) p->coef[i] = p->g_coef[0];
    for (int r = 0; r < R; r++)
      for (int ax = 0; ax < 3; ax++)
        for (uint64_t i = 0; i < n; i++) p->coef[i + n + 3 * r * n + ax * n] = p->g_coef[r + 1];
  } else {
    for (uint64_t i = 0; i < n; i++) p->coef[i] = p->g_coef[0];
    for (int r = 0; r < R; r++)
      for (int ax = 0; ax < 3; ax++)
        for (uint64_t i = 0; i < n; i++) {
          p->coef[i + n + 6 * r * n + 2 * ax * n] = p->g_coef[r + 1];
          p->coef[i + n + 6 * r * n + (2 * ax + 1) * n] = p->g_coef[r + 1];
        }
  }
}

/* utils.c:605-697 (one rank): closed-form ramp on the halo-inclusive interior,
 * BOUNDARY_SRC_VAL on the first and last x planes of U1 and U2 */
static void domain_data_fill(Params *p) {
  const int R = p->R, nnx = p->ld[0], nny = p->ld[1];
  memset(p->U1, 0, sizeof(real_t) * p->ln_domain);
  memset(p->U2, 0, sizeof(real_t) * p->ln_domain);
  if (p->U3) memset(p->U3, 0, sizeof(real_t) * p->ln_domain);
  for (int k = R; k < p->nz + R; k++)
    for (int j = R; j < p->ny + R; j++)
      for (int i = R; i < p->nx + R; i++) {
        real_t r = 1.0 / 3 * (1.0 * i / p->nx + 1.0 * j / p->ny + 1.0 * k / p->nz);
        uint64_t o = ((uint64_t)k * nny + j) * nnx + i;
        p->U1[o] = r * 1.845703;
        p->U2[o] = r * 1.845703;
        if (p->U3) p->U3[o] = r * 1.845703;
      }
  for (int k = 0; k < p->ld[2]; k++)
    for (int j = 0; j < nny; j++) {
      uint64_t b = ((uint64_t)k * nny + j) * nnx;
      p->U1[b] += BOUNDARY_SRC_VAL;
      p->U2[b] += BOUNDARY_SRC_VAL;
      p->U1[b + p->nx + 2 * R - 1] += BOUNDARY_SRC_VAL;
      p->U2[b + p->nx + 2 * R - 1] += BOUNDARY_SRC_VAL;
    }
}

static void arrays_allocate(Params *p) {
  const size_t n = p->ln_domain;
  const int nc = KINFO[p->target_kernel].ncoef;
  p->U1 = aligned_alloc(64, sizeof(real_t) * (n + 16));
  p->U2 = aligned_alloc(64, sizeof(real_t) * (n + 16));
  p->U3 = KINFO[p->target_kernel].time_order == 2 ? aligned_alloc(64, sizeof(real_t) * (n + 16)) : NULL;
  p->coef = aligned_alloc(64, sizeof(real_t) * (nc ? n * nc + 16 : 16));
  if (!p->U1 || !p->U2 || !p->coef) die("no sufficient memory");
  if (p->verbose)
    printf("[rank=0] alloc. dom(err=0):%fGiB coef(err=0):%fGiB total:%fGiB\n",
           sizeof(real_t) * 2.0 * n / (1 << 30), sizeof(real_t) * 1.0 * (nc ? n * nc : 10) / (1 << 30),
           sizeof(real_t) * (2.0 * n + (nc ? n * nc : 10)) / (1 << 30));
}

static void arrays_free(Params *p) {
  free(p->U1);
  free(p->U2);
  free(p->U3);
  free(p->coef);
}

static void check(int rc, const char *what) {
  if (rc != GR_OK) {
    fprintf(stderr, "ERROR: %s: %s\n", what, gr_last_error());
    exit(1);
  }
}

static gr_handle *gpu_setup(Params *p) {
  gr_desc d;
  memset(&d, 0, sizeof(d));
  d.kernel = p->target_kernel;
  d.dtype = GR_DTYPE;
  d.nx = p->nx;
  d.ny = p->ny;
  d.nz = p->nz;
  d.flags = p->fast ? GR_FAST : 0;
  d.t_block = 1;
  for (int i = 0; i < 11; i++) d.coef[i] = p->g_coef[i];
  gr_handle *h = NULL;
  check(gr_create(&d, &h), "gr_create");
  check(gr_upload(h, p->ld, p->U1, p->U2, p->U3, KINFO[p->target_kernel].ncoef ? p->coef : NULL), "gr_upload");
  if (p->target_ts == 3) check(gr_autotune(h, 3, NULL), "gr_autotune");
  return h;
}

/* one call of the selected time stepper (TSList[ts].func) */
static void gpu_ts(Params *p, gr_handle *h) {
  int ts = p->target_ts == 3 ? 0 : p->target_ts;
  int nt_applied = p->nt;
  check(gr_run_ts(h, ts, p->nt, p->t_dim, &nt_applied), "time stepper");
  p->nt_applied = nt_applied;
}

/* utils.c:1012-1065.  The keys and their order are girih's; values that describe
 * CPU threading or cache blocking, which the GPU stepper does not have, print as
 * girih's single-thread values (threads 1) or 0 (cache block sizes). */
static void print_param(const Params *p, const gr_stats *s) {
  printf("\n******************************************************\n");
  printf("Parameters settings\n");
  printf("******************************************************\n");
  printf("Time stepper name: %s\n", TS_NAME[p->target_ts]);
  printf("Stencil Kernel name: %s\n", KINFO[p->target_kernel].name);
  printf("Stencil Kernel semi-bandwidth: %d\n", p->R);
  printf("Stencil Kernel coefficients: %s\n", KINFO[p->target_kernel].coeff);
  printf("Precision: %s\n", DP ? "DP" : "SP");
  printf("Global domain    size:%llu    nx:%d    ny:%d    nz:%d\n", (unsigned long long)p->ln_stencils, p->nx, p->ny,
         p->nz);
  printf("Rank 0 domain    size:%llu    nx:%d    ny:%d    nz:%d\n", (unsigned long long)p->ln_stencils, p->nx, p->ny,
         p->nz);
  printf("Number of time steps: %d\n", p->nt);
  printf("Alignment size: %d Bytes\n", p->alignment);
  printf("Number of tests: %d\n", p->n_tests);
  printf("Verify:   %d\n", p->verify);
  printf("Source point enabled: %d\n", 0);
  printf("Time unroll:   %d\n", p->t_dim);
  printf("Using separate call to central line update: %d\n", 0);
  printf("Halo concatenation: %d\n", p->halo_concat);
  if (p->target_ts == 2) {
    const int tgs = p->thread_group_size > 0 ? p->thread_group_size : 1;
    printf("Enable wavefronts: %d\n", 1);
    printf("Wavefront parallel strategy: %s\n", MWD_NAME[p->mwd_type >= 0 && p->mwd_type < 4 ? p->mwd_type : 0]);
    printf("Intra-diamond width:   %d\n", (p->t_dim + 1) * 2 * p->R);
    printf("Wavefront width:  %d\n", p->t_dim * 2 * p->R + p->num_wf);
    printf("Cache block size/wf (kiB): %d\n", 0);
    printf("Total cache block size (kiB): %d\n", 0);
    printf("Next larger cache block size/wf (kiB): %d (diam_width=%d)\n", 0, (p->t_dim + 3) * 2 * p->R);
    printf("Intra-diamond prologue/epilogue MStencils: %d\n", 0);
    printf("Multi-wavefront updates: %d\n", p->num_wf);
    printf("User set thread group size: %d\n", p->thread_group_size);
    printf("Thread group size: %d\n", tgs);
    printf("Threads along z-axis: %d\n", p->thz > 0 ? p->thz : tgs);
    printf("Threads along y-axis: %d\n", p->thy > 0 ? p->thy : 1);
    printf("Threads along x-axis: %d\n", p->thx > 0 ? p->thx : 1);
    printf("Threads per cell    : %d\n", p->thc > 0 ? p->thc : 1);
    printf("Threads block: %d\n", 1);
    printf("Threads stride: %d\n", 1);
  } else {
    printf("Block size in Y: %d\n", s->tile_y);
    printf("OpenMP schedule: %s\n", "none (GPU workgroups)");
  }
  printf("GPU arithmetic: %s\n", p->fast ? "FP32 compensated (--fast)" : "girih operand order");
  printf("GPU tile: %dx%d  z-chunk: %d  fused steps: %d\n", s->tile_x, s->tile_y, s->zchunk, s->t_block);
  printf("OpenMP Threads: %d\n", 1);
  printf("Assumed usable cache size: %dKiB\n", p->cache_size);
  printf("MPI size: %d\n", 1);
  printf("Processors topology (npx, npy, npz): 01,01,01\n");
  printf("******************************************************\n");
}

/* ---- --verify: serial CPU evaluation (verification.c:52-312, stencils.c order) ---- */
#define VI(i, j, k) (((uint64_t)(k) * (uint64_t)ny + (uint64_t)(j)) * (uint64_t)nx_ + (uint64_t)(i))
static void serial_sweep(int kernel, int R, int nx_, int ny, int nz_, const real_t *coef, real_t *u, const real_t *v,
                         const real_t *roc2) {
  const uint64_t n = (uint64_t)nx_ * ny * nz_;
#define C(m, i, j, k) coef[(uint64_t)(m) * n + VI(i, j, k)]
  for (int k = R; k < nz_ - R; k++)
    for (int j = R; j < ny - R; j++)
      for (int i = R; i < nx_ - R; i++) {
        const uint64_t o = VI(i, j, k);
        real_t out;
        switch (kernel) {
          case 0: {
            real_t lap = coef[0] * v[o];
            for (int r = 1; r <= 4; r++) {
              lap = lap + coef[r] * (v[VI(i + r, j, k)] + v[VI(i - r, j, k)]);
              lap = lap + coef[r] * (v[VI(i, j + r, k)] + v[VI(i, j - r, k)]);
              lap = lap + coef[r] * (v[VI(i, j, k + r)] + v[VI(i, j, k - r)]);
            }
            out = ((real_t)2.0 * v[o] - u[o]) + roc2[o] * lap;
            break;
          }
          case 1:
            out = coef[0] * v[o] + coef[1] * (v[VI(i + 1, j, k)] + v[VI(i - 1, j, k)]) +
                  coef[1] * (v[VI(i, j + 1, k)] + v[VI(i, j - 1, k)]) + coef[1] * (v[VI(i, j, k + 1)] + v[VI(i, j, k - 1)]);
            break;
          case 2:
            out = C(0, i, j, k) * v[o] + C(1, i, j, k) * (v[VI(i + 1, j, k)] + v[VI(i - 1, j, k)]) +
                  C(1, i, j, k) * (v[VI(i, j + 1, k)] + v[VI(i, j - 1, k)]) +
                  C(1, i, j, k) * (v[VI(i, j, k + 1)] + v[VI(i, j, k - 1)]);
            break;
          case 3:
            out = C(0, i, j, k) * v[o] + C(1, i, j, k) * (v[VI(i + 1, j, k)] + v[VI(i - 1, j, k)]) +
                  C(2, i, j, k) * (v[VI(i, j + 1, k)] + v[VI(i, j - 1, k)]) +
                  C(3, i, j, k) * (v[VI(i, j, k + 1)] + v[VI(i, j, k - 1)]);
            break;
          case 4: {
            real_t l = C(0, i, j, k) * v[o];
            for (int r = 1; r <= 4; r++) {
              l = l + C(3 * r - 2, i, j, k) * (v[VI(i + r, j, k)] + v[VI(i - r, j, k)]);
              l = l + C(3 * r - 1, i, j, k) * (v[VI(i, j + r, k)] + v[VI(i, j - r, k)]);
              l = l + C(3 * r, i, j, k) * (v[VI(i, j, k + r)] + v[VI(i, j, k - r)]);
            }
            out = l;
            break;
          }
          default:
            out = C(0, i, j, k) * v[o] + C(1, i, j, k) * v[VI(i - 1, j, k)] + C(2, i, j, k) * v[VI(i + 1, j, k)] +
                  C(3, i, j, k) * v[VI(i, j - 1, k)] + C(4, i, j, k) * v[VI(i, j + 1, k)] +
                  C(5, i, j, k) * v[VI(i, j, k - 1)] + C(6, i, j, k) * v[VI(i, j, k + 1)];
        }
        u[o] = out;
      }
#undef C
}

/* The serial reference of --verify: the initial state rebuilt from scratch
 * (domain_data_fill) on unpadded copies (x extent nx+2R, as verification.c:52-312
 * allocates), with p's coefficient arrays, stepped serially for the run's
 * application count.  *pu / *pv receive U1 / U2 (caller frees both). */
static void serial_reference(const Params *p, real_t **pu, real_t **pv) {
  const int R = p->R, ux = p->nx + 2 * R, uy = p->ld[1], uz = p->ld[2];
  const uint64_t un = (uint64_t)ux * uy * uz;
  const int nc = KINFO[p->target_kernel].ncoef;
  real_t *u = calloc(un, sizeof(real_t)), *v = calloc(un, sizeof(real_t));
  real_t *r2 = p->U3 ? calloc(un, sizeof(real_t)) : NULL;
  real_t *cf = nc ? calloc(un * nc, sizeof(real_t)) : p->coef;
  if (!u || !v || (p->U3 && !r2) || !cf) die("no sufficient memory");
  Params q = *p;
  q.U1 = calloc(p->ln_domain, sizeof(real_t));
  q.U2 = calloc(p->ln_domain, sizeof(real_t));
  q.U3 = p->U3 ? calloc(p->ln_domain, sizeof(real_t)) : NULL;
  if (!q.U1 || !q.U2 || (p->U3 && !q.U3)) die("no sufficient memory");
  q.verbose = 0;
  domain_data_fill(&q);
  for (uint64_t row = 0; row < (uint64_t)uy * uz; row++) {
    memcpy(u + row * ux, q.U1 + row * p->ld[0], sizeof(real_t) * ux);
    memcpy(v + row * ux, q.U2 + row * p->ld[0], sizeof(real_t) * ux);
    if (r2) memcpy(r2 + row * ux, q.U3 + row * p->ld[0], sizeof(real_t) * ux);
    for (int m = 0; m < nc; m++)
      memcpy(cf + m * un + row * ux, p->coef + m * p->ln_domain + row * p->ld[0], sizeof(real_t) * ux);
  }
  /* applications: 2*ceil(nt/2) (ts 0/1/3) or nt-1 after rounding (ts 2); U1 is written first */
  const int napply = p->target_ts == 2 ? p->nt_applied - 1 : 2 * ((p->nt + 1) / 2);
  for (int s = 0; s < napply; s++) {
    if (s % 2 == 0)
      serial_sweep(p->target_kernel, R, ux, uy, uz, cf, u, v, r2);
    else
      serial_sweep(p->target_kernel, R, ux, uy, uz, cf, v, u, r2);
  }
  free(r2);
  if (nc) free(cf);
  free(q.U1);
  free(q.U2);
  free(q.U3);
  *pu = u;
  *pv = v;
}

static void verify(Params *p) {
  printf("#ts:%s stencil:%s|R:%d|T:%d|%s nt:%03d thrd:%d prec.:%s %s dom:(%d,%03d,%03d) top:(1,1,1) ",
         TS_NAME[p->target_ts], KINFO[p->target_kernel].name, p->R, KINFO[p->target_kernel].time_order,
         KINFO[p->target_kernel].coeff, p->nt, 1, DP ? "DP" : "SP", "no-concat", p->nx, p->ny, p->nz);
  if (p->target_ts == 2) printf("TB:%d wf:%d ", p->t_dim, 1);
  fflush(stdout);
  arrays_allocate(p);
  init_coeff(p);
  domain_data_fill(p);
  gr_handle *h = gpu_setup(p);
  gpu_ts(p, h);
  check(gr_download(h, p->ld, p->U1, p->U2), "gr_download");
  gr_destroy(h);

  real_t *u = NULL, *v = NULL;
  serial_reference(p, &u, &v);
  const int R = p->R, ux = p->nx + 2 * R, uy = p->ld[1];
  /* compare_results_std (verification.c:823-860) on U1 */
  real_t diff_l1 = 0, max_error = 0, abs_diff;
  for (int k = 0; k < p->nz; k++)
    for (int j = 0; j < p->ny; j++)
      for (int i = 0; i < p->nx; i++) {
        const uint64_t a = ((uint64_t)(k + R) * uy + (j + R)) * ux + (i + R);
        const uint64_t b = ((uint64_t)(k + R) * uy + (j + R)) * p->ld[0] + (i + R);
        abs_diff = fabs(u[a] - p->U1[b]);
        if (abs_diff > max_error) max_error = abs_diff;
        diff_l1 += abs_diff;
      }
  if ((diff_l1 > 0.0) || (diff_l1 * 0 != 0) || (diff_l1 != diff_l1)) {
    printf("Max snapshot abs. err.:%e  L1 norm:%e\n", max_error, diff_l1);
    fprintf(stderr, "BROKEN KERNEL\n");
    exit(1);
  }
  printf("eMax:%.3e|eL1:%.3e", max_error, diff_l1);
  printf("-PASSED\n");
  free(u);
  free(v);
  arrays_free(p);
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
static int cmp_d(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* performance.c:29-127 and utils.c:808-982 */
static void performance_test(Params *p) {
  time_t t0 = time(NULL);
  printf("Started on %s", ctime(&t0));
  arrays_allocate(p);
  init_coeff(p);
  domain_data_fill(p);
  gr_handle *h = gpu_setup(p); /* after autotuning (ts 3) the launch configuration is known */
  if (p->verbose) {
    gr_stats s0;
    check(gr_get_stats(h, &s0), "gr_get_stats");
    print_param(p, &s0);
  }
  printf("\n******************************************************\nPerformance results\n"
         "******************************************************\n");
  double *tt = calloc(p->n_tests > 0 ? p->n_tests : 1, sizeof(double));
  double t = 0, t_min = 1e30, t_max = -1;
  for (int i = 0; i < p->n_tests; i++) {
    const double a = now();
    gpu_ts(p, h); /* synchronous: returns when the device finished */
    const double b = now();
    tt[i] = b - a;
    t += tt[i] / p->n_tests;
    const double per = tt[i] / p->nt_applied;
    if (per < t_min) t_min = per;
    if (per > t_max) t_max = per;
    printf("Rank 0 TEST#%02d time: %e\n", i + 1, tt[i]); /* performance.c:87-95 */
    if (i == p->n_tests - 1) printf("******************************************************\n");
  }
  qsort(tt, p->n_tests, sizeof(double), cmp_d);
  const double t_med = p->n_tests > 0 ? tt[p->n_tests / 2] : 0;
  check(gr_download(h, p->ld, p->U1, p->U2), "gr_download");
  gr_stats s;
  check(gr_get_stats(h, &s), "gr_get_stats");
  gr_destroy(h);

  /* NaN / zero scan of U1 (utils.c:818-840) */
  uint64_t nans = 0, zeros = 0;
  for (uint64_t k = 0; k < p->ln_domain; k++) {
    nans += (p->U1[k] * 0) != 0;
    if (fabs(p->U1[k]) < 1e-6) zeros++;
  }
  if (100 * zeros / p->ln_domain > 90)
    printf("\n##WARNING[rank:0]: %llu%% of the sub domain contains zeroes. This might result in inaccurate performance "
           "results\n", (unsigned long long)(100 * zeros / p->ln_domain));
  if (nans)
    printf("\n##WARNING[rank:0]: %llu nan and/or -inf/inf values in the final sub domain solution. This might result in "
           "inaccurate performance results\n", (unsigned long long)nans);
  const int nt = p->nt_applied;
  printf("Total memory allocation per MPI rank: %llu MiB\n",
         (unsigned long long)(sizeof(real_t) * p->ln_domain * 3 / 1024 / 1024));
  printf("Total time(s): %e\n", t * p->n_tests);
  printf("time/test(s): %e\n", t);
  if (p->target_ts != 2) {
    printf("\nRANK0 GStencil/s MEDIAN: %f  \n", p->ln_stencils / (1e9 * t_med) * nt);
    printf("RANK0 GStencil/s    MIN: %f  \n", p->ln_stencils / (1e9 * t_max));
    printf("RANK0 GStencil/s    AVG: %f  \n", p->ln_stencils / (1e9 * (t / nt)));
    printf("RANK0 GStencil/s    MAX: %f  \n", p->ln_stencils / (1e9 * t_min));
    printf("\n******************************************************\n");
    printf("RANK0 Total: %f (s) -%06.2f%%\n", t, 100.0);
    printf("RANK0 Computation: %f (s) - %05.2f%%\n", t, 100.0);
    printf("RANK0 Communication: %f (s) - %05.2f%%\n", 0.0, 0.0);
    printf("RANK0 Waiting: %f (s) - %05.2f%%\n", 0.0, 0.0);
    printf("RANK0 Other: %f (s) - %05.2f%%\n", 0.0, 0.0);
    /* one process: the MEAN / MAX / MIN over ranks equal rank 0's figures (utils.c:884-903) */
    static const char *stat[3] = {"MEAN", "MAX", "MIN"};
    for (int q = 0; q < 3; q++) {
      printf("\n******************************************************\n");
      if (q == 0) {
        printf("MEAN Total: %f (s) -%06.2f%%\n", t, 100.0);
        printf("MEAN Computation: %f (s) - %05.2f%%\n", t, 100.0);
        printf("MEAN Communication: %f (s) - %05.2f%%\n", 0.0, 0.0);
        printf("MEAN Waiting: %f (s) - %05.2f%%\n", 0.0, 0.0);
        printf("MEAN Other: %f (s) - %05.2f%%\n", 0.0, 0.0);
      } else {
        printf("%s Total: %f (s)\n", stat[q], t);
        printf("%s Computation: %f (s)\n", stat[q], t);
        printf("%s Communication: %f (s)\n", stat[q], 0.0);
        printf("%s Waiting: %f (s)\n", stat[q], 0.0);
        printf("%s Other: %f (s)\n", stat[q], 0.0);
      }
    }
  } else {
    printf("\nTotal RANK0 MStencil/s MIN: %f  \n", p->ln_stencils / (1e6 * t_max));
    printf("Total RANK0 MStencil/s MAX: %f  \n", p->ln_stencils / (1e6 * t_min));
    printf("******************************************************\n");
    printf("MWD main-loop RANK0 MStencil/s MIN: %f\n", p->ln_stencils * (double)nt / 1e6 / (t_max * nt));
    printf("MWD main-loop RANK0 MStencil/s MAX: %f\n", p->ln_stencils * (double)nt / 1e6 / (t_min * nt));
    printf("******************************************************\n");
    /* utils.c:908-980: one GPU stands for girih's single core and thread group; the
     * whole run is the diamond main loop (no prologue, synchronisation or spin waits) */
    printf("%-27s %f (s) - %05.2f%%\n", "RANK0 ts main loop:", t, 100.0);
    printf("%-27s %f (s) - %05.2f%%\n", "RANK0 ts prologue/epilogue:", 0.0, 0.0);
    printf("%-27s %f (s) - %05.2f%%\n", "RANK0 ts others:", 0.0, 0.0);
    printf("******************************************************\n");
    printf("%-30s  core %02d     \n", "Metric \\ core:", 0);
    printf("%-27s  %e\n", "Wavefront synchronization [s]:", 0.0);
    printf("%-27s  %05.2f       \n\n", "Wavefront synchronization [%]:", 0.0);
    printf("%-27s  group %02d    \n", "Metric \\ thread group:", 0);
    printf("%-27s  %e\n", "Wavefront steady state [s]:", t);
    printf("%-27s  %05.2f       \n", "Wavefront steady state [%]:", 100.0);
    printf("%-27s  %e\n", "Wavefront startup/end [s]:", 0.0);
    printf("%-27s  %05.2f       \n", "Wavefront startup/end [%]:", 0.0);
    printf("%-27s  %e\n", "Wavefront communication [s]:", 0.0);
    printf("%-27s  %05.2f       \n", "Wavefront communication [%]:", 0.0);
    printf("%-27s  %e\n", "Wavefront others [s]:", 0.0);
    printf("%-27s  %05.2f       \n", "Wavefront others [%]:", 0.0);
    printf("%-27s  %e\n", "Group spin-wait [s]:", 0.0);
    printf("%-27s  %05.2f       \n", "Group spin-wait [%]:", 0.0);
    printf("%-27s  %e", "Resolved diamonds:", 0.0);
  }
  printf("\nGPU tile: %dx%d  z-chunk: %d  fused steps: %d  device time of last test (ms): %f\n", s.tile_x, s.tile_y,
         s.zchunk, s.t_block, s.last_ms);
  printf("\n******************************************************\n");
  free(tt);
  arrays_free(p);
  time_t t1 = time(NULL);
  printf("COMPLETED SUCCESSFULLY on %s", ctime(&t1));
}

int main(int argc, char **argv) {
  Params p;
  param_default(&p);
  parse_args(argc, argv, &p);
  init(&p);
  if (p.verify)
    verify(&p);
  else
    performance_test(&p);
  return 0;
}
