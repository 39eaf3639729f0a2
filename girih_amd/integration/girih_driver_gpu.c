/*
 * girih_driver_gpu.c -- girih's mwd_kernel with the MI355X stepper registered.
 *
 * A maintainer makes the GPU selectable by adding one TSList entry
 * (src/wrappers.h:28-36); this file is that edit made in a translation unit of
 * our own, so that girih's sources stay untouched: it defines TSList with girih's
 * three CPU steppers plus "MI355X" (gpu_ts, index 3), and then follows girih's
 * driver flow (src/driver.c:21-82): MPI_Init_thread -> param_default ->
 * parse_args -> mpi_topology_init -> init -> verify | performance_test.  Linked
 * with girih's own objects (utils, mpi_utils, performance, verification, the
 * kernels), every other line that runs is girih's: argument parsing, the
 * closed-form initial state, the serial verification and its exact compare, the
 * timing protocol and the report.
 *
 * Built with -DGPU_AS_DIAMOND, index 2 ("Diamond") runs gpu_diamond_ts instead of
 * girih's CPU diamond stepper, so girih's diamond initialisation (validity rules
 * and nt rounding, intra_diamond_info_init) precedes the GPU run.
 *
 * For ts 3, girih's init() has no case (source information, utils.c:391-399) and
 * mpi_halo_init has none either (mpi_utils.c:325-333); with one process and the
 * source point disabled (girih's default, utils.c:336) neither is needed.
 */
#include <stdio.h>

#include "data_structures.h"

extern void naive_nonblocking_ts(Parameters *p);
extern void halo_first_ts(Parameters *p);
extern void dynamic_intra_diamond_ts(Parameters *p);
extern void gpu_ts(Parameters *p);
extern void gpu_diamond_ts(Parameters *p);

extern void performance_test(Parameters *p);
extern void param_default(Parameters *p);
extern void parse_args(int argc, char **argv, Parameters *p);
extern void mpi_topology_init(Parameters *p);
extern void init(Parameters *p);
extern void verify(Parameters *p);

struct time_stepper TSList[] = {
    {"Spatial Blocking", naive_nonblocking_ts},
    {"Halo-first", halo_first_ts},
#ifdef GPU_AS_DIAMOND
    {"Diamond", gpu_diamond_ts},
#else
    {"Diamond", dynamic_intra_diamond_ts},
#endif
    {"MI355X", gpu_ts},
    {0, 0},
};

int main(int argc, char **argv) {
  int provided;
  Parameters p;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_MULTIPLE, &provided);
  MPI_Comm_rank(MPI_COMM_WORLD, &p.mpi_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &p.mpi_size);
  param_default(&p);
  parse_args(argc, argv, &p);
  if (p.t.shape[0] * p.t.shape[1] * p.t.shape[2] != p.mpi_size) {
    if (p.mpi_rank == 0) fprintf(stderr, "ERROR: requested MPI topology shape does not match the process count\n");
    MPI_Finalize();
    return 1;
  }
  mpi_topology_init(&p);
  init(&p);
  if (p.verify != 0)
    verify(&p);
  else
    performance_test(&p);
  MPI_Finalize();
  return 0;
}
