/*
 * girih_oracle.h -- CPU oracle for the star-stencil time stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of girih's
 * serial verification reference (src/verification.c) used to check the
 * HIP stepper.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never calls it.
 *
 * Arrays use the unpadded halo-inclusive layout of verification.c:61-64:
 *   shape = {nx+2R, ny+2R, nz+2R},  idx(i,j,k) = (k*shape[1] + j)*shape[0] + i
 * Variable coefficients are SoA: coef[m*n_domain + idx]   (verification.c COEF macro).
 *
 * Kernel ids follow girih's stencil_info_list (src/kernels/stencils.c:260-269):
 *   0 25-pt star, 2nd order in time, constant coef   (iso_ref)
 *   1  7-pt star, 1st order, constant coef           (iso_ref_2space_1time)
 *   2  7-pt star, 1st order, variable (2 arrays)     (iso_ref_2space_1time_var)
 *   3  7-pt star, 1st order, var axis-sym (4 arrays) (iso_ref_2space_1time_var_axsym)
 *   4 25-pt star, 1st order, var axis-sym (13 arr.)  (iso_ref_8space_1time_var_axsym)
 *   5  7-pt star, 1st order, var no-sym (7 arrays)   (iso_ref_2space_1time_var_nosym)
 */
#ifndef GIRIH_ORACLE_H_
#define GIRIH_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stencil radius, time order and number of coefficient arrays (0 = 5 scalars) */
int gro_radius(int kernel);
int gro_time_order(int kernel);
int gro_num_coef_arrays(int kernel);

/* girih's default coefficient list g_coef (utils.c:116-119) rounded to the type */
void gro_default_coef_f64(double out[11]);
void gro_default_coef_f32(float out[11]);

/* nt as the reference's diamond stepper rounds it (diamond_utils.c:1042-1056) */
int gro_mwd_round_nt(int nt, int t_dim);

/* One sweep over the interior [R, shape-R) of every axis (verification.c:315-479). */
int gro_sweep_f64(int kernel, const int shape[3], const double *coef,
                  double *u, const double *v, const double *roc2);
int gro_sweep_f32(int kernel, const int shape[3], const float *coef,
                  float *u, const float *v, const float *roc2);

/* Closed-form initial state of verification.c:219-251 (u, v, roc2 may be NULL). */
void gro_init_ref_f64(int kernel, int nx, int ny, int nz, double *u, double *v, double *roc2);
void gro_init_ref_f32(int kernel, int nx, int ny, int nz, float *u, float *v, float *roc2);

/* Coefficients of verification.c:92-216 built from g_coef (n_domain = prod(shape)). */
void gro_init_coef_f64(int kernel, uint64_t n_domain, const double g_coef[11], double *coef);
void gro_init_coef_f32(int kernel, uint64_t n_domain, const float g_coef[11], float *coef);

/*
 * Time loop of verification.c:281-284: for(it=0; it<nt; it+=2){ u<-f(v); v<-f(u); }.
 * When src_idx >= 0 the point source of nb_naive_ts.c:190,196 is injected:
 * u[src_idx] += src[it] after the first half step, v[src_idx] += src[it+1]
 * after the second (src must hold 2*ceil(nt/2) values).
 * Returns the number of applied steps (2*ceil(nt/2)), or -1 on a bad kernel id.
 */
int gro_run_f64(int kernel, const int shape[3], const double *coef, double *u, double *v,
                const double *roc2, int nt, int64_t src_idx, const double *src);
int gro_run_f32(int kernel, const int shape[3], const float *coef, float *u, float *v,
                const float *roc2, int nt, int64_t src_idx, const float *src);

/* Exact comparison of compare_results_std (verification.c:823-860): the interior of the
 * unpadded oracle array u against target (nx*ny*nz, no halo).  Writes max |diff| and
 * the L1 sum of |diff| accumulated in the array type; returns 1 when diff_l1 == 0. */
int gro_compare_f64(const double *u, const double *target, int nx, int ny, int nz, int R,
                    double *emax, double *el1);
int gro_compare_f32(const float *u, const float *target, int nx, int ny, int nz, int R,
                    double *emax, double *el1);

/*
 * FP32 kernel 0 in the library's GR_FAST arithmetic (no girih counterpart): per-radius
 * sums (radii 1-2 as differences from the centre) scaled with fmaf from radius 4 to 1,
 * centre weight (float)(c0 + 6*(c1 + c2)) rounded once from double, increment-form
 * update u = v + fmaf(roc2, lap, v - u).  roc2 == NULL uses the scalar roc2s.
 * gro_run_fast_f32 is gro_run_f32's time loop (with its source) on this sweep.
 */
void gro_sweep_fast_f32(const int shape[3], const double coef[5], float *u, const float *v,
                        const float *roc2, double roc2s);
int gro_run_fast_f32(const int shape[3], const double coef[5], float *u, float *v, const float *roc2,
                     double roc2s, int nt, int64_t src_idx, const float *src);

/*
 * Seismic shot on kernel 0 in FP64 (adjoint.c; no girih counterpart): forward
 * traces of plane rec_k (trace [nt][ny][nx], may be NULL) and, when res and grad
 * are given, the adjoint gradient of sum_a <res[a], trace[a]> with respect to the
 * squared slowness m = dt_over_h^2 / roc2 (grad [nz][ny][nx]).  Zero initial
 * fields; src_idx < 0 for no source.  Returns 0, or < 0 on bad arguments.
 */
int gro_shot_f64(const int shape[3], const double coef[5], const double *roc2, int nt, int64_t src_idx,
                 const double *src, int rec_k, double *trace, const double *res, double *grad,
                 double dt_over_h);

#ifdef __cplusplus
}
#endif
#endif
