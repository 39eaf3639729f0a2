/*
 * gpu_ts.c -- the MI355X stepper as a girih time stepper.
 *
 * This is the file a girih maintainer adds to girih's build to make the GPU a
 * selectable target: it is compiled against girih's own headers
 * (src/data_structures.h: Parameters, real_t, Profile) and linked with
 * libgirih_gpu.so (include/girih_gpu.h).  Each function has the time-stepper
 * signature void (*)(Parameters *) of struct time_stepper
 * (src/data_structures.h:294-297) and goes into TSList (src/wrappers.h:28-36):
 *
 *   gpu_ts          spatial-blocking semantics (2*ceil(nt/2) applications; girih's
 *                   naive_nonblocking_ts, src/kernels/nb_naive_ts.c:187-203) -- the
 *                   new "MI355X" entry;
 *   gpu_diamond_ts  the diamond stepper's semantics (nt as girih's init rounded it,
 *                   nt-1 applications; dynamic_intra_diamond_ts,
 *                   src/kernels/diamond_ts.c:871-976) -- girih's "Diamond" entry run
 *                   on the GPU, so that girih's own diamond initialisation
 *                   (intra_diamond_info_init) precedes it.
 *
 * Contract of a girih stepper (verification.c:37, performance.c:75): read
 * p->U1/U2/U3/coef in girih's padded layout, leave the results in p->U1 and
 * p->U2, accumulate p->prof.  The first call creates one device problem for the
 * process (girih runs one problem per process), uploads the host arrays and
 * autotunes the launch; every call then steps the device state and copies U1 and
 * U2 back, so repeated calls (girih's n_tests) continue from the last state like
 * the CPU steppers do.  A failing library call ends the run the way girih's
 * RAISE_ERROR does (data_structures.h:303-314).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "data_structures.h"
#include "girih_gpu.h"

static gr_handle *g_gpu = NULL;

static void gpu_destroy(void) {
  gr_destroy(g_gpu);
  g_gpu = NULL;
}

/* RAISE_ERROR's behaviour with a message that is not a format string */
static void gpu_fail(Parameters *p, const char *what) {
  if (p->mpi_rank == 0) fprintf(stderr, "ERROR: %s: %s\n", what, gr_last_error());
  MPI_Barrier(MPI_COMM_WORLD);
  MPI_Finalize();
  exit(1);
}

static void gpu_setup(Parameters *p) {
  gr_desc d;
  int i;
  if (p->mpi_size != 1) {
    if (p->mpi_rank == 0) fprintf(stderr, "ERROR: the MI355X stepper runs one process on one GPU\n");
    MPI_Barrier(MPI_COMM_WORLD);
    MPI_Finalize();
    exit(1);
  }
  memset(&d, 0, sizeof(d));
  d.kernel = p->target_kernel;
  d.dtype = sizeof(real_t) == 8 ? GR_F64 : GR_F32;
  d.nx = p->lstencil_shape[0];
  d.ny = p->lstencil_shape[1];
  d.nz = p->lstencil_shape[2];
  d.t_block = 0; /* single sweeps until the autotuner picks a fused depth */
  /* constant-coefficient kernels: girih's coef holds coef[0..R] (utils.c:436-496) */
  if (p->stencil.coeff == CONSTANT_COEFFICIENT)
    for (i = 0; i <= p->stencil.r && i < 11; i++) d.coef[i] = p->coef[i];
  if (gr_create(&d, &g_gpu) != GR_OK) gpu_fail(p, "gr_create");
  atexit(gpu_destroy);
  if (gr_upload(g_gpu, p->ldomain_shape, p->U1, p->U2, p->stencil.time_order == 2 ? p->U3 : NULL,
                p->stencil.coeff == CONSTANT_COEFFICIENT ? NULL : p->coef) != GR_OK)
    gpu_fail(p, "gr_upload");
  if (gr_autotune(g_gpu, 3, NULL) != GR_OK) gpu_fail(p, "gr_autotune");
}

static void gpu_run(Parameters *p, int semantics) {
  int nt_applied = 0;
  double t0, t1;
  if (!g_gpu) gpu_setup(p);
  t0 = MPI_Wtime();
  if (gr_run_ts(g_gpu, semantics, p->nt, p->t_dim, &nt_applied) != GR_OK) gpu_fail(p, "gr_run_ts");
  if (gr_download(g_gpu, p->ldomain_shape, p->U1, p->U2) != GR_OK) gpu_fail(p, "gr_download");
  t1 = MPI_Wtime();
  p->prof.compute += t1 - t0;
  if (semantics == 2) p->prof.ts_main += t1 - t0;
}

void gpu_ts(Parameters *p) { gpu_run(p, 0); }

void gpu_diamond_ts(Parameters *p) { gpu_run(p, 2); }
