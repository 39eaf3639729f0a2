/*
 * adjoint.c -- FP64 CPU reference for seismic shots on the 25-point wave stepper.
 *
 * TEST INFRASTRUCTURE ONLY (see girih_oracle.h).  girih has no adjoint; this file
 * restates the mathematics the GPU shot code implements, on top of the oracle's
 * kernel-0 sweep (verification.c:315-352 operand order), keeping every state in
 * memory (small grids only):
 *
 *   forward   x_0 = x_1 = 0,
 *             x_{a+2} = 2 x_{a+1} - x_a + roc2 * L x_{a+1}   (gro_sweep_f64, kernel 0)
 *             x_{a+2}[src] += src[a]                          (nb_naive_ts.c:190,196)
 *             trace[a][j][i] = x_{a+2}(i+R, j+R, rec_k)
 *   adjoint   nu_{nt+2} = nu_{nt+3} = 0,
 *             nu_{a+2} = 2 nu_{a+3} - nu_{a+4} + roc2 * L nu_{a+3}   (the same sweep)
 *             nu_{a+2}(plane rec_k) += roc2 * res[a]
 *             g += nu_{a+2} * ((x_{a+2} - x_{a+1}) - (x_{a+1} - x_a) - src[a] e_src)
 *   gradient  dJ/dm = -g / dt_over_h^2,   m = 1/v^2, roc2 = dt_over_h^2 / m.
 *
 * nu = roc2 * mu, mu the Lagrange multiplier of x, which turns the adjoint of the
 * leapfrog (whose spatial operator is L^T diag(roc2) = L diag(roc2), L symmetric on
 * the interior with girih's zero halo) into the forward recurrence itself.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "girih_oracle.h"

int gro_shot_f64(const int shape[3], const double coef[5], const double *roc2, int nt, int64_t src_idx,
                 const double *src, int rec_k, double *trace, const double *res, double *grad,
                 double dt_over_h) {
  const int R = 4;
  const int nnx = shape[0], nny = shape[1], nnz = shape[2];
  const int nx = nnx - 2 * R, ny = nny - 2 * R, nz = nnz - 2 * R;
  const size_t n = (size_t)nnx * nny * nnz;
  double *x, *n0, *n1, *g, *tmp;
  int a, i, j, k;
  if (nt < 1 || nx < 1 || ny < 1 || nz < 1 || rec_k < R || rec_k >= nnz - R) return -1;
  x = (double *)calloc((size_t)(nt + 2) * n, sizeof(double));
  if (!x) return -2;
  for (a = 0; a < nt; a++) {
    double *xa = x + (size_t)a * n, *x1 = x + (size_t)(a + 1) * n, *x2 = x + (size_t)(a + 2) * n;
    memcpy(x2, xa, n * sizeof(double)); /* the sweep's u holds x_a on entry */
    gro_sweep_f64(0, shape, coef, x2, x1, roc2);
    if (src_idx >= 0) x2[src_idx] += src[a];
    if (trace)
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++)
          trace[((size_t)a * ny + j) * nx + i] = x2[((size_t)rec_k * nny + j + R) * nnx + i + R];
  }
  if (res && grad) {
    n0 = (double *)calloc(n, sizeof(double));
    n1 = (double *)calloc(n, sizeof(double));
    g = (double *)calloc((size_t)nx * ny * nz, sizeof(double));
    if (!n0 || !n1 || !g) {
      free(x);
      free(n0);
      free(n1);
      free(g);
      return -2;
    }
    for (a = nt - 1; a >= 0; a--) {
      const double *xa = x + (size_t)a * n, *x1 = x + (size_t)(a + 1) * n, *x2 = x + (size_t)(a + 2) * n;
      gro_sweep_f64(0, shape, coef, n0, n1, roc2); /* n0 <- nu_{a+2} */
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
          const size_t o = ((size_t)rec_k * nny + j + R) * nnx + i + R;
          n0[o] = n0[o] + roc2[o] * res[((size_t)a * ny + j) * nx + i];
        }
      for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
          for (i = 0; i < nx; i++) {
            const size_t o = ((size_t)(k + R) * nny + j + R) * nnx + i + R;
            double d = (x2[o] - x1[o]) - (x1[o] - xa[o]);
            if (src_idx >= 0 && (int64_t)o == src_idx) d = d - src[a];
            g[((size_t)k * ny + j) * nx + i] += n0[o] * d;
          }
      tmp = n0;
      n0 = n1;
      n1 = tmp;
    }
    {
      const double scale = -1.0 / (dt_over_h * dt_over_h);
      for (i = 0; i < nx * ny * nz; i++) grad[i] = scale * g[i];
    }
    free(n0);
    free(n1);
    free(g);
  }
  free(x);
  return 0;
}
