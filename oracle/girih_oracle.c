/*
 * girih_oracle.c -- CPU oracle for the star-stencil time stepper.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of girih's serial verification
 * reference (src/verification.c:52-479, 823-860), the initial state and
 * coefficients it builds, and the step pairing of src/kernels/nb_naive_ts.c:187-203.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Build: gcc -O0 -ffp-contract=off (see oracle/Makefile).  The reference compiles
 * its oracle at -O0 (Makefile:4,51-52); x86-64 baseline code has no FMA, so the
 * restatement reproduces the reference bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include "girih_oracle.h"

/* stencil_info_list (src/kernels/stencils.c:260-269) for the regular star kernels */
static const int k_radius[6] = {4, 1, 1, 1, 4, 1};
static const int k_time_order[6] = {2, 1, 1, 1, 1, 1};
/* coefficient arrays per kernel: utils.c:183-201 (0 = constant scalars) */
static const int k_ncoef[6] = {0, 0, 2, 4, 13, 7};

int gro_radius(int kernel) { return (kernel >= 0 && kernel < 6) ? k_radius[kernel] : -1; }
int gro_time_order(int kernel) { return (kernel >= 0 && kernel < 6) ? k_time_order[kernel] : -1; }
int gro_num_coef_arrays(int kernel) { return (kernel >= 0 && kernel < 6) ? k_ncoef[kernel] : -1; }

/* utils.c:116-119 */
static const double g_coef_default[11] = {-0.28472, 0.16000, -0.02000, 0.00254, -0.00018,
                                          -0.18472, 0.19,    -0.0500,  0.00554, -0.0009,
                                          0.00354};
void gro_default_coef_f64(double out[11]) {
  int i;
  for (i = 0; i < 11; i++) out[i] = g_coef_default[i];
}
void gro_default_coef_f32(float out[11]) {
  int i;
  for (i = 0; i < 11; i++) out[i] = (float)g_coef_default[i];
}

/* diamond_utils.c:1042-1056 (regular stencils) */
int gro_mwd_round_nt(int nt, int t_dim) {
  int w = (t_dim + 1) * 2;
  int remain = (nt - 2) % w;
  if (remain != 0) nt = nt + w - remain;
  return nt;
}

#define REAL double
#define SFX(name) name##_f64
#define TWO 2.
#define FABS fabs
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef TWO
#undef FABS

#define REAL float
#define SFX(name) name##_f32
#define TWO 2.f
#define FABS fabsf
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef TWO
#undef FABS
