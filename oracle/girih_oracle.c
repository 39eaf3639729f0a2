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

/*
 * FP32 kernel 0 in the library's GR_FAST arithmetic (no girih counterpart; girih's
 * FP32 build uses the operand order of k0_wave25 above).  Radii 1-2 are summed as
 * differences from the centre, radii 3-4 as pair sums, each radius scaled with one
 * correctly rounded fused multiply-add, from radius 4 down to 1; the centre weight
 * c0 + 6*(c1 + c2) is rounded once from double; the update is the increment form
 *   u = v + fma(roc2, lap, v - u).
 * The oracle keeps this as its own function so that tests can compare the GPU's
 * GR_FAST kernels bit for bit; the function itself is checked against the FP64
 * reference order (k0_wave25_f64) within a stated tolerance and against the exact
 * Laplacian of a quadratic field.  roc2 == NULL uses the scalar roc2s.
 */
void gro_sweep_fast_f32(const int shape[3], const double coef[5], float *u, const float *v,
                        const float *roc2, double roc2s) {
  const int nnx = shape[0], nny = shape[1], nnz = shape[2];
  const uint64_t sy = (uint64_t)nnx, sz = (uint64_t)nnx * nny;
  float c[5];
  double s = coef[0];
  int r, i, j, k;
  for (r = 0; r < 5; r++) c[r] = (float)coef[r];
  s += 6.0 * coef[1];
  s += 6.0 * coef[2];
  const float csum = (float)s, rs = (float)roc2s;
  for (k = 4; k < nnz - 4; k++)
    for (j = 4; j < nny - 4; j++)
      for (i = 4; i < nnx - 4; i++) {
        const uint64_t o = (uint64_t)k * sz + (uint64_t)j * sy + (uint64_t)i;
        const float v0 = v[o];
        float l = csum * v0;
        for (r = 4; r >= 1; r--) {
          const float xp = v[o + r], xm = v[o - r];
          const float yp = v[o + r * sy], ym = v[o - r * sy];
          const float zp = v[o + r * sz], zm = v[o - r * sz];
          float d;
          if (r <= 2)
            d = (((xp - v0) + (xm - v0)) + ((yp - v0) + (ym - v0))) + ((zp - v0) + (zm - v0));
          else
            d = ((xp + xm) + (yp + ym)) + (zp + zm);
          l = fmaf(c[r], d, l);
        }
        const float rc = roc2 ? roc2[o] : rs;
        u[o] = v0 + fmaf(rc, l, v0 - u[o]);
      }
}

/* The time loop of gro_run_f32 on gro_sweep_fast_f32 (same source injection). */
int gro_run_fast_f32(const int shape[3], const double coef[5], float *u, float *v, const float *roc2,
                     double roc2s, int nt, int64_t src_idx, const float *src) {
  int it;
  for (it = 0; it < nt; it += 2) {
    gro_sweep_fast_f32(shape, coef, u, v, roc2, roc2s);
    if (src_idx >= 0 && src) u[src_idx] += src[it];
    gro_sweep_fast_f32(shape, coef, v, u, roc2, roc2s);
    if (src_idx >= 0 && src) v[src_idx] += src[it + 1];
  }
  return 2 * ((nt + 1) / 2);
}
