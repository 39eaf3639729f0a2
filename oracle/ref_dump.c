/*
 * ref_dump.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A main() that replaces girih's src/driver.c so that the reference's own stepper
 * output can be captured.  It is linked against the reference objects compiled
 * from /root/reference/src where they lie (oracle/Makefile); nothing of the
 * reference is copied.  Flow (mirrors verification.c:26-51):
 *   param_default -> parse_args -> mpi_topology_init -> init -> arrays_allocate
 *   -> init_coeff -> domain_data_fill -> TSList[ts].func -> dump U1,U2
 *   -> aggregate_subdomains -> verify_serial_generic (prints eMax..-PASSED or exits 1)
 *
 * usage: ref_dump_{dp,sp} OUT.bin <mwd_kernel flags...>
 * OUT.bin: int32 header[12] = {magic 0x47524846, sizeof(real_t), nx, ny, nz, R,
 *          nnx, nny, nnz (padded ldomain_shape), nt (as applied), ts, kernel}
 *          followed by U1 then U2 (nnx*nny*nnz values each, reference layout).
 */
#include <stdio.h>
#include <stdlib.h>
#include "driver.h"
#include "verification.h"

extern void mpi_halo_init(Parameters *);
extern void mpi_halo_finalize(Parameters *);

int main(int argc, char **argv) {
  int provided;
  if (argc < 2) {
    fprintf(stderr, "usage: %s OUT.bin [mwd_kernel flags]\n", argv[0]);
    return 2;
  }
  MPI_Init_thread(&argc, &argv, MPI_THREAD_MULTIPLE, &provided);
  Parameters p;
  MPI_Comm_rank(MPI_COMM_WORLD, &p.mpi_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &p.mpi_size);
  param_default(&p);
  parse_args(argc - 1, argv + 1, &p); /* argv[1] (the output path) stands in for argv[0] */
  mpi_topology_init(&p);
  init(&p);

  mpi_halo_init(&p);
  arrays_allocate(&p);
  init_coeff(&p);
  domain_data_fill(&p);
  TSList[p.target_ts].func(&p);

  FILE *f = fopen(argv[1], "wb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    MPI_Abort(MPI_COMM_WORLD, 2);
  }
  int hdr[12] = {0x47524846, (int)sizeof(real_t), p.stencil_shape[0], p.stencil_shape[1],
                 p.stencil_shape[2], p.stencil.r, p.ldomain_shape[0], p.ldomain_shape[1],
                 p.ldomain_shape[2], p.nt, p.target_ts, p.target_kernel};
  fwrite(hdr, sizeof(int), 12, f);
  fwrite(p.U1, sizeof(real_t), p.ln_domain, f);
  fwrite(p.U2, sizeof(real_t), p.ln_domain, f);
  fclose(f);

  real_t *aggr = aggregate_subdomains(p);
  verify_serial_generic(aggr, p);
  free(aggr);
  arrays_free(&p);
  mpi_halo_finalize(&p);
  MPI_Finalize();
  return 0;
}
