/*
 * oracle.c -- CPU restatement of the vtkmakeCompare time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed beside it.  The product never links or calls this file.
 *
 * Each function cites the reference routine it restates and keeps its
 * operation order (compile with -ffp-contract=off so that no multiply-add
 * is fused).  Loops that the reference runs under OpenMP carry the same
 * "omp parallel for"; the reference's serial passes stay serial.
 *
 * Pinning (see tests/): every line solver, stencil, boundary routine,
 * set_levelset and the two adjoint convection kernels are compared
 * bit-for-bit (or to 1e-14 for complex division) against the reference
 * Fortran compiled from its own sources (oracle/_ref); the FFT against
 * numpy.fft and single-mode spikes; MKP against a brute-force sparse solve
 * and the post-projection divergence; set_adj_NS_scalar against closed
 * forms (its reference reads an uninitialised local and cannot be run).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FIELD(a, i, j, k) a[(size_t)(i) + sy * (size_t)(j) + sz * (size_t)(k)]
#define CVAL(i, j, k, n)                                                      \
  cellvalue[(size_t)((i)-1) +                                                 \
            (size_t)nx * ((size_t)((j)-1) +                                   \
                          (size_t)ny * ((size_t)((k)-1) +                     \
                                        (size_t)nz * (size_t)((n)-1)))]
#define GEOM                                                                   \
  const size_t sy = (size_t)nx + 2;                                           \
  const size_t sz = ((size_t)nx + 2) * ((size_t)ny + 2);                      \
  (void)sy;                                                                   \
  (void)sz

/* ---------------------------------------------------------------- grid */

/* nse_dominik.f:553-607 */
void o_set_c(int nx, int ny, int nz, double *x, double *x12, double *y,
             double *y12, double *z, double *z12, double *dx, double *dx12,
             double *dy, double *dy12, double *dz, double *dz12, double xl,
             double yl, double zl) {
  for (int i = 0; i <= nx + 1; i++) {
    x12[i] = xl * (double)i / (double)nx;
    x[i] = xl * ((double)i - 0.50) / (double)nx;
  }
  for (int j = 0; j <= ny + 1; j++) {
    y12[j] = yl * (double)j / (double)ny;
    y[j] = yl * ((double)j - 0.50) / (double)ny;
  }
  for (int k = 0; k <= nz + 1; k++) {
    z12[k] = zl * (double)k / (double)nz;
    z[k] = zl * ((double)k - 0.50) / (double)nz;
  }
  for (int i = 0; i <= nx; i++) dx12[i] = x[i + 1] - x[i];
  dx12[nx + 1] = 2.0 * (x12[nx + 1] - x[nx + 1]);
  for (int i = 1; i <= nx + 1; i++) dx[i] = x12[i] - x12[i - 1];
  dx[0] = dx[1];
  dx[nx + 1] = dx[nx];
  for (int j = 0; j <= ny; j++) dy12[j] = y[j + 1] - y[j];
  dy12[ny + 1] = 2.0 * (y12[ny + 1] - y[ny + 1]);
  for (int j = 1; j <= ny + 1; j++) dy[j] = y12[j] - y12[j - 1];
  dy[0] = dy[1];
  dy[ny + 1] = dy[ny];
  for (int k = 0; k <= nz; k++) dz12[k] = z[k + 1] - z[k];
  dz12[nz + 1] = 2.0 * (z12[nz + 1] - z[nz + 1]);
  for (int k = 1; k <= nz + 1; k++) dz[k] = z12[k] - z12[k - 1];
  dz[0] = dz[1];
  dz[nz + 1] = dz[nz];
}

/* ------------------------------------------------------------- set_bc */

/* nse_dominik.f:611-753: u,w,T reflect at the walls, v is zero on the wall
   faces j=0 and j=ny, p is Neumann; x and z periodic.  The order x, y, z
   per field fixes the corner values. */
void o_set_bc(double *u, double *v, double *w, double *p, int nx, int ny,
              int nz, double *t) {
  GEOM;
  for (int which = 0; which < 5; which++) {
    double *a;
    int ymode; /* 0 reflect, 1 zero faces 0 and ny, 2 Neumann */
    switch (which) {
      case 0: a = u; ymode = 0; break;
      case 1: a = v; ymode = 1; break;
      case 2: a = w; ymode = 0; break;
      case 3: a = p; ymode = 2; break;
      default: a = t; ymode = 0; break;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz + 1; k++)
      for (int j = 0; j <= ny + 1; j++) {
        FIELD(a, 0, j, k) = FIELD(a, nx, j, k);
        FIELD(a, nx + 1, j, k) = FIELD(a, 1, j, k);
      }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz + 1; k++)
      for (int i = 0; i <= nx + 1; i++) {
        if (ymode == 0) {
          FIELD(a, i, 0, k) = -1.0 * FIELD(a, i, 1, k);
          FIELD(a, i, ny + 1, k) = -1.0 * FIELD(a, i, ny, k);
        } else if (ymode == 1) {
          FIELD(a, i, 0, k) = 0.0;
          FIELD(a, i, ny, k) = 0.0;
        } else {
          FIELD(a, i, 0, k) = FIELD(a, i, 1, k);
          FIELD(a, i, ny + 1, k) = FIELD(a, i, ny, k);
        }
      }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) {
        FIELD(a, i, j, 0) = FIELD(a, i, j, nz);
        FIELD(a, i, j, nz + 1) = FIELD(a, i, j, 1);
      }
  }
}

/* ------------------------------------------------------------ stencils */

/* nse_dominik.f:757-831 */
void o_set_convect(const double *u, double *u_v, const double *v, double *v_v,
                   const double *w, double *w_v, double dt, int nx, int ny,
                   int nz, const double *dx, const double *dy, const double *dz,
                   const double *dx12, const double *dy12,
                   const double *dz12) {
  GEOM;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(u, i, j, k);
        double u1 = ((c + FIELD(u, i + 1, j, k)) * (c + FIELD(u, i + 1, j, k)) -
                     (c + FIELD(u, i - 1, j, k)) * (c + FIELD(u, i - 1, j, k))) /
                    dx12[i];
        double u2 = ((FIELD(v, i + 1, j, k) * dx[i + 1] + FIELD(v, i, j, k) * dx[i]) *
                         (FIELD(u, i, j + 1, k) + c) -
                     (FIELD(v, i + 1, j - 1, k) * dx[i + 1] +
                      FIELD(v, i, j - 1, k) * dx[i]) *
                         (c + FIELD(u, i, j - 1, k))) /
                    dx12[i] / dy[j];
        double u3 = ((FIELD(w, i + 1, j, k) * dx[i + 1] + FIELD(w, i, j, k) * dx[i]) *
                         (FIELD(u, i, j, k + 1) + c) -
                     (FIELD(w, i + 1, j, k - 1) * dx[i + 1] +
                      FIELD(w, i, j, k - 1) * dx[i]) *
                         (c + FIELD(u, i, j, k - 1))) /
                    dx12[i] / dz[k];
        FIELD(u_v, i, j, k) = 0.25 * (-u1 - u2 - u3) * dt + 1.0 * dt;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(v, i, j, k);
        double v1 = ((FIELD(u, i, j + 1, k) * dy[j + 1] + FIELD(u, i, j, k) * dy[j]) *
                         (FIELD(v, i + 1, j, k) + c) -
                     (FIELD(u, i - 1, j + 1, k) * dy[j + 1] +
                      FIELD(u, i - 1, j, k) * dy[j]) *
                         (c + FIELD(v, i - 1, j, k))) /
                    dy12[j] / dx[i];
        double v2 = ((c + FIELD(v, i, j + 1, k)) * (c + FIELD(v, i, j + 1, k)) -
                     (c + FIELD(v, i, j - 1, k)) * (c + FIELD(v, i, j - 1, k))) /
                    dy12[j];
        double v3 = ((FIELD(w, i, j + 1, k) * dy[j + 1] + FIELD(w, i, j, k) * dy[j]) *
                         (FIELD(v, i, j, k + 1) + c) -
                     (FIELD(w, i, j + 1, k - 1) * dy[j + 1] +
                      FIELD(w, i, j, k - 1) * dy[j]) *
                         (c + FIELD(v, i, j, k - 1))) /
                    dy12[j] / dz[k];
        FIELD(v_v, i, j, k) = 0.25 * (-v1 - v2 - v3) * dt;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(w, i, j, k);
        double w1 = ((FIELD(u, i, j, k + 1) * dz[k + 1] + FIELD(u, i, j, k) * dz[k]) *
                         (FIELD(w, i + 1, j, k) + c) -
                     (FIELD(u, i - 1, j, k + 1) * dz[k + 1] +
                      FIELD(u, i - 1, j, k) * dz[k]) *
                         (c + FIELD(w, i - 1, j, k))) /
                    dz12[k] / dx[i];
        double w2 = ((FIELD(v, i, j, k + 1) * dz[k + 1] + FIELD(v, i, j, k) * dz[k]) *
                         (FIELD(w, i, j + 1, k) + c) -
                     (FIELD(v, i, j - 1, k + 1) * dz[k + 1] +
                      FIELD(v, i, j - 1, k) * dz[k]) *
                         (c + FIELD(w, i, j - 1, k))) /
                    dz12[k] / dy[j];
        double w3 = ((c + FIELD(w, i, j, k + 1)) * (c + FIELD(w, i, j, k + 1)) -
                     (c + FIELD(w, i, j, k - 1)) * (c + FIELD(w, i, j, k - 1))) /
                    dz12[k];
        FIELD(w_v, i, j, k) = 0.25 * (-w1 - w2 - w3) * dt;
      }
}

/* nse_dominik.f:834-884 */
void o_set_convection_for_scalar(const double *u, const double *v,
                                 const double *w, double dt, int nx, int ny,
                                 int nz, const double *dx, const double *dy,
                                 const double *dz, const double *dx12,
                                 const double *dy12, const double *dz12,
                                 const double *t, double *u_vt,
                                 const double *rhocp) {
  GEOM;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double rc = FIELD(rhocp, i, j, k), tc = FIELD(t, i, j, k);
        double ut = (-((FIELD(rhocp, i - 1, j, k) * FIELD(t, i - 1, j, k) * dx[i] +
                        rc * tc * dx[i - 1]) *
                       FIELD(u, i - 1, j, k)) /
                         2.0 / dx12[i - 1] +
                     ((rc * tc * dx[i + 1] +
                       FIELD(rhocp, i + 1, j, k) * FIELD(t, i + 1, j, k) * dx[i]) *
                      FIELD(u, i, j, k)) /
                         2.0 / dx12[i]) /
                    dx[i];
        double vt = (-((FIELD(rhocp, i, j - 1, k) * FIELD(t, i, j - 1, k) * dy[j] +
                        rc * tc * dy[j - 1]) *
                       FIELD(v, i, j - 1, k)) /
                         2.0 / dy12[j - 1] +
                     ((rc * tc * dy[j + 1] +
                       FIELD(rhocp, i, j + 1, k) * FIELD(t, i, j + 1, k) * dy[j]) *
                      FIELD(v, i, j, k)) /
                         2.0 / dy12[j]) /
                    dy[j];
        double wt = (-((FIELD(rhocp, i, j, k - 1) * FIELD(t, i, j, k - 1) * dz[k] +
                        rc * tc * dz[k - 1]) *
                       FIELD(w, i, j, k - 1)) /
                         2.0 / dz12[k - 1] +
                     ((rc * tc * dz[k + 1] +
                       FIELD(rhocp, i, j, k + 1) * FIELD(t, i, j, k + 1) * dz[k]) *
                      FIELD(w, i, j, k)) /
                         2.0 / dz12[k]) /
                    dz[k];
        FIELD(u_vt, i, j, k) = (-ut - vt - wt) * dt / rc + 1.0 * dt;
      }
}

/* nse_dominik.f:888-945 */
void o_set_conduction(const double *u, double *u_d, const double *v,
                      double *v_d, const double *w, double *w_d, double re,
                      double dt, int nx, int ny, int nz, const double *dx,
                      const double *dy, const double *dz, const double *dx12,
                      const double *dy12, const double *dz12) {
  GEOM;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(u, i, j, k);
        double u1 = ((FIELD(u, i + 1, j, k) - c) / dx[i + 1] -
                     (c - FIELD(u, i - 1, j, k)) / dx[i]) /
                    dx12[i];
        double u2 = ((FIELD(u, i, j + 1, k) - c) / dy12[j] -
                     (c - FIELD(u, i, j - 1, k)) / dy12[j - 1]) /
                    dy[j];
        double u3 = ((FIELD(u, i, j, k + 1) - c) / dz12[k] -
                     (c - FIELD(u, i, j, k - 1)) / dz12[k - 1]) /
                    dz[k];
        FIELD(u_d, i, j, k) = dt / re / 2.0 * (u1 + u2 + u3);
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(v, i, j, k);
        double v1 = ((FIELD(v, i + 1, j, k) - c) / dx12[i] -
                     (c - FIELD(v, i - 1, j, k)) / dx12[i - 1]) /
                    dx[i];
        double v2 = ((FIELD(v, i, j + 1, k) - c) / dy[j + 1] -
                     (c - FIELD(v, i, j - 1, k)) / dy[j]) /
                    dy12[j];
        double v3 = ((FIELD(v, i, j, k + 1) - c) / dz12[k] -
                     (c - FIELD(v, i, j, k - 1)) / dz12[k - 1]) /
                    dz[k];
        FIELD(v_d, i, j, k) = dt / re / 2.0 * (v1 + v2 + v3);
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double c = FIELD(w, i, j, k);
        double w1 = ((FIELD(w, i + 1, j, k) - c) / dx12[i] -
                     (c - FIELD(w, i - 1, j, k)) / dx12[i - 1]) /
                    dx[i];
        double w2 = ((FIELD(w, i, j + 1, k) - c) / dy12[j] -
                     (c - FIELD(w, i, j - 1, k)) / dy12[j - 1]) /
                    dy[j];
        double w3 = ((FIELD(w, i, j, k + 1) - c) / dz[k + 1] -
                     (c - FIELD(w, i, j, k - 1)) / dz[k]) /
                    dz12[k];
        FIELD(w_d, i, j, k) = dt / re / 2.0 * (w1 + w2 + w3);
      }
}

/* harmonic-mean face conductance (d1+d2)/(d1/k1+d2/k2) used by
   nse_dominik.f:969-997 and 1753-1769 */
static inline double hmean(double d1, double k1, double d2, double k2) {
  return (d1 + d2) / (d1 / k1 + d2 / k2);
}

/* nse_dominik.f:950-1004 */
void o_set_diffusion_for_scalar(double re, double dt, int nx, int ny, int nz,
                                const double *dx, const double *dy,
                                const double *dz, const double *dx12,
                                const double *dy12, const double *dz12,
                                const double *t, double *u_dt, double sc,
                                const double *thc, const double *rhocp) {
  GEOM;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double tc = FIELD(t, i, j, k), kc = FIELD(thc, i, j, k);
        double ut1 = (hmean(dx[i + 1], FIELD(thc, i + 1, j, k), dx[i], kc) *
                          (FIELD(t, i + 1, j, k) - tc) / dx12[i] -
                      hmean(dx[i - 1], FIELD(thc, i - 1, j, k), dx[i], kc) *
                          (tc - FIELD(t, i - 1, j, k)) / dx12[i - 1]) /
                     dx[i];
        double ut2 = (hmean(dy[j + 1], FIELD(thc, i, j + 1, k), dy[j], kc) *
                          (FIELD(t, i, j + 1, k) - tc) / dy12[j] -
                      hmean(dy[j - 1], FIELD(thc, i, j - 1, k), dy[j], kc) *
                          (tc - FIELD(t, i, j - 1, k)) / dy12[j - 1]) /
                     dy[j];
        double ut3 = (hmean(dz[k + 1], FIELD(thc, i, j, k + 1), dz[k], kc) *
                          (FIELD(t, i, j, k + 1) - tc) / dz12[k] -
                      hmean(dz[k - 1], FIELD(thc, i, j, k - 1), dz[k], kc) *
                          (tc - FIELD(t, i, j, k - 1)) / dz12[k - 1]) /
                     dz[k];
        FIELD(u_dt, i, j, k) =
            (ut1 + ut2 + ut3) / 2.0 / re / sc / FIELD(rhocp, i, j, k) * dt;
      }
}

/* nse_dominik.f:1011-1129 */
void o_set_immersed_boundary(double eta, int nx, int ny, int nz, double dt,
                             double *u_l, double *v_l, double *w_l,
                             const double *dx, const double *dy,
                             const double *dz, double *t_l, double *cellvalue,
                             const double *u_ta, const double *v_ta,
                             const double *w_ta, int l, double *cons_cell) {
  GEOM;
  (void)dx; (void)dy; (void)dz;
  if (l == 1) {
    /* PHI_INT (:1028-1035) is computed but never used */
    *cons_cell = 0.0;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx - 1; i++)
          CVAL(i, j, k, 2) = (CVAL(i, j, k, 1) + CVAL(i + 1, j, k, 1)) * 0.50;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) CVAL(nx, j, k, 2) = CVAL(nx, j, k, 1);
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny - 1; j++)
          CVAL(i, j, k, 3) = (CVAL(i, j, k, 1) + CVAL(i, j + 1, k, 1)) * 0.50;
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++) CVAL(i, ny, k, 3) = CVAL(i, ny, k, 1);
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        for (int k = 1; k <= nz - 1; k++)
          CVAL(i, j, k, 4) = (CVAL(i, j, k, 1) + CVAL(i, j, k + 1, 1)) * 0.50;
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) CVAL(i, j, nz, 4) = CVAL(i, j, nz, 1);
  }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(u_l, i, j, k) = eta * CVAL(i, j, k, 2) * FIELD(u_ta, i, j, k) * dt;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(v_l, i, j, k) = eta * CVAL(i, j, k, 3) * FIELD(v_ta, i, j, k) * dt;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(w_l, i, j, k) = eta * CVAL(i, j, k, 4) * FIELD(w_ta, i, j, k) * dt;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) FIELD(t_l, i, j, k) = 0.0;
}

/* AB2 weights of nse_dominik.f:1152-1160 */
static void ab2(int l, double *a1, double *a2, double *b1) {
  if (l == 1) {
    *a1 = 1.0; *a2 = 0.0; *b1 = 1.0;
  } else {
    *a1 = 1.5; *a2 = -0.5; *b1 = 1.0;
  }
}

/* one velocity component of nse_dominik.f:1165-1197 */
static void explicit_component(const double *xv, const double *xpre,
                               const double *xd, double *xa, const double *xl,
                               int nx, int ny, int nz, double a1, double a2,
                               double b1) {
  GEOM;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(xa, i, j, k) = a1 * FIELD(xv, i, j, k) + a2 * FIELD(xpre, i, j, k) +
                             b1 * FIELD(xd, i, j, k) + b1 * FIELD(xl, i, j, k);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      FIELD(xa, 0, j, k) = FIELD(xa, nx, j, k);
      FIELD(xa, nx + 1, j, k) = FIELD(xa, 1, j, k);
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      FIELD(xa, i, 0, k) = -1.0 * FIELD(xa, i, 1, k);
      FIELD(xa, i, ny + 1, k) = -1.0 * FIELD(xa, i, ny, k);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      FIELD(xa, i, j, 0) = FIELD(xa, i, j, nz);
      FIELD(xa, i, j, nz + 1) = FIELD(xa, i, j, 1);
    }
}

/* nse_dominik.f:1132-1286 */
void o_set_explicit(const double *u_v, double *u_pre, const double *u_d,
                    double *u_a, const double *v_v, double *v_pre,
                    const double *v_d, double *v_a, const double *w_v,
                    double *w_pre, const double *w_d, double *w_a,
                    const double *u_l, const double *v_l, const double *w_l,
                    int nx, int ny, int nz, int l) {
  GEOM;
  double a1, a2, b1;
  ab2(l, &a1, &a2, &b1);
  explicit_component(u_v, u_pre, u_d, u_a, u_l, nx, ny, nz, a1, a2, b1);
  explicit_component(v_v, v_pre, v_d, v_a, v_l, nx, ny, nz, a1, a2, b1);
  explicit_component(w_v, w_pre, w_d, w_a, w_l, nx, ny, nz, a1, a2, b1);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        FIELD(u_pre, i, j, k) = FIELD(u_v, i, j, k);
        FIELD(v_pre, i, j, k) = FIELD(v_v, i, j, k);
        FIELD(w_pre, i, j, k) = FIELD(w_v, i, j, k);
      }
}

/* nse_dominik.f:1291-1341, including the z halo line :1328 that copies the
   freshly written u_T(i,j,0) */
void o_set_explicit_for_tempreture(const double *u_vt, const double *u_dt,
                                   double *u_t, double *t_pre,
                                   const double *t_l, int nx, int ny, int nz,
                                   int l) {
  GEOM;
  double a1, a2, b1;
  ab2(l, &a1, &a2, &b1);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(u_t, i, j, k) = a1 * FIELD(u_vt, i, j, k) + a2 * FIELD(t_pre, i, j, k) +
                              b1 * FIELD(u_dt, i, j, k) + b1 * FIELD(t_l, i, j, k);
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++) {
      FIELD(u_t, 0, j, k) = FIELD(u_t, nx, j, k);
      FIELD(u_t, nx + 1, j, k) = FIELD(u_t, 1, j, k);
    }
  for (int i = 0; i <= nx + 1; i++)
    for (int j = 0; j <= ny + 1; j++) {
      FIELD(u_t, i, j, 0) = FIELD(u_t, i, j, nz);
      FIELD(u_t, i, j, nz + 1) = FIELD(u_t, i, j, 0);
    }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) FIELD(t_pre, i, j, k) = FIELD(u_vt, i, j, k);
}

/* nse_dominik.f:387-396 */
void o_t_divide(double *u_t, const double *t, const double *cellvalue,
                double eta, double dt, int nx, int ny, int nz) {
  GEOM;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(u_t, i, j, k) = (FIELD(t, i, j, k) + FIELD(u_t, i, j, k)) /
                              (1.0 + eta * CVAL(i, j, k, 1) * dt);
}

/* ------------------------------------------------------- line solvers */

/* Backward-then-forward elimination of nse_dominik.f:1549-1570 on a
   tridiagonal matrix whose rows 0..m are (lo[r], di[r], up[r]) for
   columns (r-1, r, r+1).  The reference stores the matrix densely, but
   every operation it performs touches only these three diagonals, so this
   is the same arithmetic in O(m) storage. */
static void ul_eliminate(double *lo, double *di, double *up, double *u,
                         int m) {
  for (int j = m; j >= 1; j--) {
    double aa = di[j];
    di[j] = di[j] / aa;
    lo[j] = lo[j] / aa;
    double bb = up[j - 1];
    up[j - 1] = up[j - 1] - bb * di[j];
    di[j - 1] = di[j - 1] - bb * lo[j];
    u[j] = u[j] / aa;
    u[j - 1] = u[j - 1] - bb * u[j];
  }
  for (int j = 0; j <= m - 1; j++) {
    double aa = di[j];
    di[j] = di[j] / aa;
    double bb = lo[j + 1];
    lo[j + 1] = lo[j + 1] - bb * di[j];
    u[j] = u[j] / aa;
    u[j + 1] = u[j + 1] - bb * u[j];
  }
}

/* nse_dominik.f:1505-1577 */
void o_implicit_uy(double *suy, int ny, const double *dy, const double *dy12,
                   double re, double dt, double ft, double fb, double eta,
                   const double *cuy) {
  int m = ny + 1;
  double *lo = calloc((size_t)4 * (m + 1), sizeof(double));
  double *di = lo + (m + 1), *up = di + (m + 1), *u = up + (m + 1);
  double beta = dt / re / 2.0;
  for (int j = 1; j <= ny; j++) {
    double alpha = dt * eta * cuy[j - 1] + 1.0;
    lo[j] = -beta / alpha / dy12[j - 1] / dy[j];
    di[j] = 1.0 + beta / alpha * (1.0 / dy12[j] + 1.0 / dy12[j - 1]) / dy[j];
    up[j] = -beta / alpha / dy12[j] / dy[j];
  }
  di[0] = 1.0; up[0] = 1.0;
  di[ny + 1] = 1.0; lo[ny + 1] = 1.0;
  for (int j = 1; j <= ny; j++) u[j] = suy[j];
  u[0] = ft;
  u[ny + 1] = fb;
  ul_eliminate(lo, di, up, u, m);
  for (int j = 0; j <= ny + 1; j++) suy[j] = u[j];
  free(lo);
}

/* nse_dominik.f:1583-1648 */
void o_implicit_vy(double *svy, int ny, const double *dy, const double *dy12,
                   double re, double dt, double ft, double fb, double eta,
                   const double *cvy) {
  int m = ny;
  double *lo = calloc((size_t)4 * (m + 1), sizeof(double));
  double *di = lo + (m + 1), *up = di + (m + 1), *u = up + (m + 1);
  double beta = dt / re / 2.0;
  di[0] = 1.0;
  di[ny] = 1.0;
  for (int j = 1; j <= ny - 1; j++) {
    double alpha = dt * eta * cvy[j - 1] + 1.0;
    lo[j] = -beta / alpha / dy[j] / dy12[j];
    di[j] = 1.0 + beta / alpha * (1.0 / dy[j + 1] + 1.0 / dy[j]) / dy12[j];
    up[j] = -beta / alpha / dy[j + 1] / dy12[j];
  }
  for (int j = 1; j <= ny - 1; j++) u[j] = svy[j];
  u[0] = ft;
  u[ny] = fb;
  ul_eliminate(lo, di, up, u, m);
  for (int j = 0; j <= ny; j++) svy[j] = u[j];
  free(lo);
}

/* periodic_implicit.f:47-64 (Numerical Recipes tridag on rows 0..n+1) */
void o_tdma(const double *a, const double *b, const double *c, const double *r,
            double *u, int n) {
  double *gam = malloc(sizeof(double) * (size_t)(n + 2));
  double bet = b[0];
  u[0] = r[0] / bet;
  for (int j = 1; j <= n + 1; j++) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    u[j] = (r[j] - a[j] * u[j - 1]) / bet;
  }
  for (int j = n; j >= 0; j--) u[j] = u[j] - gam[j + 1] * u[j + 1];
  free(gam);
}

/* periodic_implicit.f:67-93 (Sherman-Morrison with gamma = 1) */
void o_cyclic(const double *a, const double *b, const double *c, double alpha,
              double beta, const double *r, double *x, int n) {
  double *bb = malloc(sizeof(double) * (size_t)3 * (n + 2));
  double *u = bb + (n + 2), *z = u + (n + 2);
  double gamma = 1.0;
  bb[0] = b[0] - gamma;
  bb[n + 1] = b[n + 1] - alpha * beta / gamma;
  for (int i = 1; i <= n; i++) bb[i] = b[i];
  o_tdma(a, bb, c, r, x, n);
  u[0] = gamma;
  u[n + 1] = alpha;
  for (int i = 1; i <= n; i++) u[i] = 0.0;
  o_tdma(a, bb, c, u, z, n);
  double fact = (x[0] + beta * x[n + 1] / gamma) /
                (1.0 + z[0] + beta * z[n + 1] / gamma);
  for (int i = 0; i <= n + 1; i++) x[i] = x[i] - fact * z[i];
  free(bb);
}

/* periodic ghost rows of periodic_implicit.f:34-41: x0 = xn, x(n+1) = x1 */
static void periodic_ghost_rows(double *aa, double *bb, double *cc, int n) {
  aa[0] = 0.0; bb[0] = 0.0; cc[0] = 1.0;
  aa[n + 1] = 1.0; bb[n + 1] = 0.0; cc[n + 1] = 0.0;
}

/* periodic_implicit.f:1-44; cux[i-1] holds CUX(i) */
void o_implicit_solver(double *sux, int nx, const double *dx,
                       const double *dx12, double re, double dt, double ft,
                       double fb, double eta, const double *cux) {
  double *buf = calloc((size_t)4 * (nx + 2), sizeof(double));
  double *u = buf, *aa = buf + (nx + 2), *bb = aa + (nx + 2), *cc = bb + (nx + 2);
  double beta = dt / re / 2.0;
  for (int i = 1; i <= nx; i++) u[i] = sux[i];
  u[0] = ft;
  u[nx + 1] = fb;
  for (int i = 1; i <= nx; i++) {
    double alpha = dt * eta * cux[i - 1] + 1.0;
    aa[i] = -beta / alpha / dx[i] / dx12[i];
    bb[i] = 1.0 + beta / alpha * (1.0 / dx[i + 1] + 1.0 / dx[i]) / dx12[i];
    cc[i] = -beta / alpha / dx[i + 1] / dx12[i];
  }
  periodic_ghost_rows(aa, bb, cc, nx);
  o_cyclic(aa, bb, cc, -1.0, -1.0, u, sux, nx);
  free(buf);
}

/* nse_dominik.f:1725-1781; cux is (0:nx+1) as declared there */
void o_implicit_tx(double *stx, int nx, const double *dx, const double *dx12,
                   double re, double sc, double dt, const double *llx,
                   const double *rrx, double ft, double fb, double eta,
                   const double *cux) {
  double *buf = calloc((size_t)4 * (nx + 2), sizeof(double));
  double *tt = buf, *aa = buf + (nx + 2), *bb = aa + (nx + 2), *cc = bb + (nx + 2);
  double beta = dt / re / sc / 2.0;
  for (int i = 1; i <= nx; i++) tt[i] = stx[i];
  tt[0] = ft;
  tt[nx + 1] = fb;
  for (int i = 1; i <= nx; i++) {
    double alpha = dt * eta * cux[i] + 1.0;
    double h1 = hmean(dx[i], llx[i], dx[i - 1], llx[i - 1]);
    double h2 = hmean(dx[i + 1], llx[i + 1], dx[i], llx[i]);
    aa[i] = -beta / alpha * h1 / dx12[i - 1] / dx[i] / rrx[i];
    bb[i] = 1.0 + beta / alpha * (h2 / dx12[i] / dx[i] + h1 / dx12[i - 1] / dx[i]) /
                      rrx[i];
    cc[i] = -beta / alpha * h2 / dx12[i] / dx[i] / rrx[i];
  }
  periodic_ghost_rows(aa, bb, cc, nx);
  o_cyclic(aa, bb, cc, -1.0, -1.0, tt, stx, nx);
  free(buf);
}

/* nse_dominik.f:1784-1870; cuy is (0:ny+1) as declared there */
void o_implicit_ty(double *sty, int ny, const double *dy, const double *dy12,
                   double re, double sc, double dt, const double *lly,
                   const double *rry, double ft, double fb, const double *cuy,
                   double eta) {
  int m = ny + 1;
  double *lo = calloc((size_t)4 * (m + 1), sizeof(double));
  double *di = lo + (m + 1), *up = di + (m + 1), *tt = up + (m + 1);
  double beta = dt / re / sc / 2.0;
  di[0] = 1.0; up[0] = 1.0;
  di[ny + 1] = 1.0; lo[ny + 1] = 1.0;
  for (int i = 1; i <= ny; i++) {
    double alpha = dt * eta * cuy[i] + 1.0;
    double h1 = hmean(dy[i], lly[i], dy[i - 1], lly[i - 1]);
    double h2 = hmean(dy[i + 1], lly[i + 1], dy[i], lly[i]);
    lo[i] = -beta / alpha * h1 / dy12[i - 1] / dy[i] / rry[i];
    di[i] = 1.0 + beta / alpha * (h2 / dy12[i] / dy[i] + h1 / dy12[i - 1] / dy[i]) /
                      rry[i];
    up[i] = -beta / alpha * h2 / dy12[i] / dy[i] / rry[i];
  }
  for (int i = 1; i <= ny; i++) tt[i] = sty[i];
  tt[0] = ft;
  tt[ny + 1] = fb;
  ul_eliminate(lo, di, up, tt, m);
  for (int i = 1; i <= ny; i++) sty[i] = tt[i];
  free(lo);
}

/* nse_dominik.f:1344-1461 with the gathers parallel_uy/vy/z/x
   (nse_dominik.f:2893-2967) */
void o_solve_u_temp(double *u_a, double *v_a, double *w_a, int nx, int ny,
                    int nz, const double *dx, const double *dy,
                    const double *dz, const double *dx12, const double *dy12,
                    const double *dz12, double re, double dt, double eta,
                    const double *cellvalue, const double *u, const double *v,
                    const double *w) {
  GEOM;
  double *comp[3] = {u_a, v_a, w_a};
  const double *old[3] = {u, v, w};
  for (int c = 0; c < 3; c++) {
    double *fa = comp[c];
    const double *fo = old[c];
    int n = c + 2; /* cellvalue slot 2,3,4 */
    int jmax = (c == 1) ? ny - 1 : ny; /* v's z and x sweeps stop at ny-1 */
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++)
          FIELD(fa, i, j, k) = (FIELD(fo, i, j, k) + FIELD(fa, i, j, k)) /
                               (1.0 + eta * CVAL(i, j, k, n) * dt);
    /* y sweep */
#pragma omp parallel
    {
      double *s = malloc(sizeof(double) * (size_t)(ny + 2));
      double *cv = malloc(sizeof(double) * (size_t)(ny + 2));
#pragma omp for schedule(static)
      for (int i = 1; i <= nx; i++)
        for (int k = 1; k <= nz; k++) {
          if (c == 1) {
            for (int j = 1; j <= ny - 1; j++) {
              s[j] = FIELD(fa, i, j, k);
              cv[j - 1] = CVAL(i, j, k, n);
            }
            o_implicit_vy(s, ny, dy, dy12, re, dt, 0.0, 0.0, eta, cv);
            for (int j = 0; j <= ny; j++) FIELD(fa, i, j, k) = s[j];
          } else {
            for (int j = 1; j <= ny; j++) {
              s[j] = FIELD(fa, i, j, k);
              cv[j - 1] = CVAL(i, j, k, n);
            }
            o_implicit_uy(s, ny, dy, dy12, re, dt, 0.0, 0.0, eta, cv);
            for (int j = 0; j <= ny + 1; j++) FIELD(fa, i, j, k) = s[j];
          }
        }
      free(s);
      free(cv);
    }
    /* z sweep */
#pragma omp parallel
    {
      double *s = malloc(sizeof(double) * (size_t)(nz + 2));
      double *cv = malloc(sizeof(double) * (size_t)(nz + 2));
#pragma omp for schedule(static)
      for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= jmax; j++) {
          for (int k = 1; k <= nz; k++) {
            s[k] = FIELD(fa, i, j, k);
            cv[k - 1] = CVAL(i, j, k, n);
          }
          o_implicit_solver(s, nz, dz, dz12, re, dt, 0.0, 0.0, eta, cv);
          for (int k = 0; k <= nz + 1; k++) FIELD(fa, i, j, k) = s[k];
        }
      free(s);
      free(cv);
    }
    /* x sweep */
#pragma omp parallel
    {
      double *s = malloc(sizeof(double) * (size_t)(nx + 2));
      double *cv = malloc(sizeof(double) * (size_t)(nx + 2));
#pragma omp for schedule(static)
      for (int j = 1; j <= jmax; j++)
        for (int k = 1; k <= nz; k++) {
          for (int i = 1; i <= nx; i++) {
            s[i] = FIELD(fa, i, j, k);
            cv[i - 1] = CVAL(i, j, k, n);
          }
          o_implicit_solver(s, nx, dx, dx12, re, dt, 0.0, 0.0, eta, cv);
          for (int i = 0; i <= nx + 1; i++) FIELD(fa, i, j, k) = s[i];
        }
      free(s);
      free(cv);
    }
  }
}

/* nse_dominik.f:1465-1501 with parallel_Ty/Tx/Tz (nse_dominik.f:2969-3045).
   The conductivity ghosts the reference leaves unset are defined here:
   x and z wrap periodically, y copies the wall-adjacent cell. */
void o_solve_t(double *u_t, int nx, int ny, int nz, const double *dx,
               const double *dy, const double *dz, const double *dx12,
               const double *dy12, const double *dz12, double re, double sc,
               double dt, const double *rhocp, const double *thc,
               const double *cellvalue, double eta) {
  GEOM;
  int nmax = nx > ny ? nx : ny;
  nmax = nmax > nz ? nmax : nz;
#pragma omp parallel
  {
    double *s = calloc((size_t)4 * (nmax + 2), sizeof(double));
    double *cu = s + (nmax + 2), *rr = cu + (nmax + 2), *ll = rr + (nmax + 2);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; i++)
      for (int k = 1; k <= nz; k++) {
        for (int j = 1; j <= ny; j++) {
          rr[j] = FIELD(rhocp, i, j, k);
          ll[j] = FIELD(thc, i, j, k);
          s[j] = FIELD(u_t, i, j, k);
          cu[j] = CVAL(i, j, k, 1);
        }
        ll[0] = ll[1];
        ll[ny + 1] = ll[ny];
        o_implicit_ty(s, ny, dy, dy12, re, sc, dt, ll, rr, 0.0, 0.0, cu, eta);
        for (int j = 1; j <= ny; j++) FIELD(u_t, i, j, k) = s[j];
      }
    free(s);
  }
#pragma omp parallel
  {
    double *s = calloc((size_t)4 * (nmax + 2), sizeof(double));
    double *cu = s + (nmax + 2), *rr = cu + (nmax + 2), *ll = rr + (nmax + 2);
#pragma omp for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int k = 1; k <= nz; k++) {
        for (int i = 1; i <= nx; i++) {
          rr[i] = FIELD(rhocp, i, j, k);
          ll[i] = FIELD(thc, i, j, k);
          s[i] = FIELD(u_t, i, j, k);
          cu[i] = CVAL(i, j, k, 1);
        }
        ll[0] = ll[nx];
        ll[nx + 1] = ll[1];
        o_implicit_tx(s, nx, dx, dx12, re, sc, dt, ll, rr, 0.0, 0.0, eta, cu);
        for (int i = 1; i <= nx; i++) FIELD(u_t, i, j, k) = s[i];
      }
    free(s);
  }
#pragma omp parallel
  {
    double *s = calloc((size_t)4 * (nmax + 2), sizeof(double));
    double *cu = s + (nmax + 2), *rr = cu + (nmax + 2), *ll = rr + (nmax + 2);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; i++)
      for (int j = 1; j <= ny; j++) {
        for (int k = 1; k <= nz; k++) {
          rr[k] = FIELD(rhocp, i, j, k);
          ll[k] = FIELD(thc, i, j, k);
          s[k] = FIELD(u_t, i, j, k);
          cu[k] = CVAL(i, j, k, 1);
        }
        ll[0] = ll[nz];
        ll[nz + 1] = ll[1];
        o_implicit_tx(s, nz, dz, dz12, re, sc, dt, ll, rr, 0.0, 0.0, eta, cu);
        for (int k = 1; k <= nz; k++) FIELD(u_t, i, j, k) = s[k];
      }
    free(s);
  }
}

/* -------------------------------------------------------------- FFT */

/* Recursive mixed-radix (2, 3, generic) decimation-in-time complex DFT,
   out of place into `out` (contiguous), sign -1 forward.  Twiddles come
   from one table tw[m] = exp(sign 2 pi i m / n0) of the top-level size n0
   (libm cos/sin); accuracy is checked against numpy.fft in tests/. */
static void fft_rec(const double *in, int n, int stride, double *out,
                    double *work, const double *tw, int n0) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : n;
  const int step = n0 / n;
  if (p == n) { /* direct DFT */
    for (int k = 0; k < n; k++) {
      double sr = 0.0, si = 0.0;
      for (int m = 0; m < n; m++) {
        long e = ((long)k * m) % n * step;
        double c = tw[2 * e], s = tw[2 * e + 1];
        double xr = in[2 * (size_t)m * stride], xi = in[2 * (size_t)m * stride + 1];
        sr += xr * c - xi * s;
        si += xr * s + xi * c;
      }
      out[2 * k] = sr;
      out[2 * k + 1] = si;
    }
    return;
  }
  int m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + 2 * (size_t)q * stride, m, stride * p, work + 2 * (size_t)q * m, out, tw,
            n0);
  for (int k = 0; k < m; k++) {
    double tr[3], ti[3];
    for (int q = 0; q < p; q++) {
      long e = (long)q * k * step;
      double c = tw[2 * e], s = tw[2 * e + 1];
      double xr = work[2 * ((size_t)q * m + k)], xi = work[2 * ((size_t)q * m + k) + 1];
      tr[q] = xr * c - xi * s;
      ti[q] = xr * s + xi * c;
    }
    for (int r = 0; r < p; r++) {
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < p; q++) {
        long e = (long)((q * r) % p) * (n0 / p);
        double c = tw[2 * e], s = tw[2 * e + 1];
        sr += tr[q] * c - ti[q] * s;
        si += tr[q] * s + ti[q] * c;
      }
      out[2 * ((size_t)r * m + k)] = sr;
      out[2 * ((size_t)r * m + k) + 1] = si;
    }
  }
}

void o_fft(double *data, int n, int stride, int sign) {
  double *buf = malloc(sizeof(double) * 8 * (size_t)n);
  double *out = buf, *work = buf + 2 * (size_t)n, *in = work + 2 * (size_t)n;
  double *tw = in + 2 * (size_t)n;
  for (int m = 0; m < n; m++) {
    double ang = sign * 2.0 * M_PI * (double)m / (double)n;
    tw[2 * m] = cos(ang);
    tw[2 * m + 1] = sin(ang);
    in[2 * m] = data[2 * (size_t)m * stride];
    in[2 * m + 1] = data[2 * (size_t)m * stride + 1];
  }
  fft_rec(in, n, 1, out, work, tw, n);
  for (int m = 0; m < n; m++) {
    data[2 * (size_t)m * stride] = out[2 * m];
    data[2 * (size_t)m * stride + 1] = out[2 * m + 1];
  }
  free(buf);
}

/* p2s_fftw2d.f:5-114.  Per plane J: FI(I+1,K+1) = FP(I,J,K) for
   I,K = 0..N-1 (index 0 is the periodic image of N), a 2D real-to-complex
   DFT of size NX x NZ keeping I = 0..NX/2, then division by NX*NZ. */
void o_p2s(const double *fp, double *fpt, int nx, int ny, int nz) {
  GEOM;
  const int nh = nx / 2;
  const size_t ty = (size_t)nh + 1, tz = ((size_t)nh + 1) * ((size_t)ny + 2);
#pragma omp parallel
  {
    double *plane = malloc(sizeof(double) * 2 * (size_t)nx * nz);
#pragma omp for schedule(static)
    for (int j = 1; j <= ny; j++) {
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
          plane[2 * ((size_t)i + (size_t)nx * k)] = FIELD(fp, i, j, k);
          plane[2 * ((size_t)i + (size_t)nx * k) + 1] = 0.0;
        }
      for (int k = 0; k < nz; k++) o_fft(plane + 2 * (size_t)nx * k, nx, 1, -1);
      for (int i = 0; i <= nh; i++) o_fft(plane + 2 * (size_t)i, nz, nx, -1);
      double scale = (double)(nx * nz);
      for (int k = 0; k < nz; k++)
        for (int i = 0; i <= nh; i++) {
          size_t o = 2 * ((size_t)i + ty * j + tz * k);
          fpt[o] = plane[2 * ((size_t)i + (size_t)nx * k)] / scale;
          fpt[o + 1] = plane[2 * ((size_t)i + (size_t)nx * k) + 1] / scale;
        }
    }
    free(plane);
  }
}

/* p2s_fftw2d.f:116-223.  Unnormalised complex-to-real inverse: the
   half spectrum I = 0..NX/2 is completed by Hermitian symmetry (the
   imaginary parts of the I = 0 and I = NX/2 columns act only through that
   completion, as in a c2r transform), then FP(I,J,K) for I,K = 0..N-1 and
   the wrap FP(NX,J,K) = FP(0,J,K), FP(I,J,NZ) = FP(I,J,0). */
void o_s2p(const double *fpt, double *fp, int nx, int ny, int nz) {
  GEOM;
  const int nh = nx / 2;
  const size_t ty = (size_t)nh + 1, tz = ((size_t)nh + 1) * ((size_t)ny + 2);
#pragma omp parallel
  {
    double *plane = malloc(sizeof(double) * 2 * (size_t)nx * nz);
#pragma omp for schedule(static)
    for (int j = 1; j <= ny; j++) {
      for (int k = 0; k < nz; k++)
        for (int i = 0; i <= nh; i++) {
          size_t o = 2 * ((size_t)i + ty * j + tz * k);
          plane[2 * ((size_t)i + (size_t)nx * k)] = fpt[o];
          plane[2 * ((size_t)i + (size_t)nx * k) + 1] = fpt[o + 1];
        }
      /* inverse along z for the stored columns */
      for (int i = 0; i <= nh; i++) o_fft(plane + 2 * (size_t)i, nz, nx, +1);
      /* complete each x row by Hermitian symmetry and invert along x */
      for (int k = 0; k < nz; k++) {
        double *row = plane + 2 * (size_t)nx * k;
        row[1] = 0.0;
        if (nx % 2 == 0) row[2 * nh + 1] = 0.0;
        for (int i = nh + 1; i < nx; i++) {
          row[2 * i] = row[2 * (nx - i)];
          row[2 * i + 1] = -row[2 * (nx - i) + 1];
        }
        o_fft(row, nx, 1, +1);
      }
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++)
          FIELD(fp, i, j, k) = plane[2 * ((size_t)i + (size_t)nx * k)];
      for (int k = 0; k <= nz; k++) FIELD(fp, nx, j, k) = FIELD(fp, 0, j, k);
      for (int i = 0; i <= nx; i++) FIELD(fp, i, j, nz) = FIELD(fp, i, j, 0);
    }
    free(plane);
  }
}

/* ----------------------------------------------------- spectral solves */

/* tridiag.f:43-58.  The complex/real divisions are done per component
   (the Fortran mixed-mode division may differ in the last bit). */
void o_ctridiag(double *a, double *b, int nn) {
  double *a1 = a, *a2 = a + nn, *a3 = a + 2 * (size_t)nn;
  for (int i = 0; i < nn - 1; i++) {
    a3[i] = a3[i] / a2[i];
    b[2 * i] = b[2 * i] / a2[i];
    b[2 * i + 1] = b[2 * i + 1] / a2[i];
    a2[i + 1] = a2[i + 1] - a1[i + 1] * a3[i];
    b[2 * (i + 1)] = b[2 * (i + 1)] - a1[i + 1] * b[2 * i];
    b[2 * (i + 1) + 1] = b[2 * (i + 1) + 1] - a1[i + 1] * b[2 * i + 1];
  }
  b[2 * (nn - 1)] = b[2 * (nn - 1)] / a2[nn - 1];
  b[2 * (nn - 1) + 1] = b[2 * (nn - 1) + 1] / a2[nn - 1];
  for (int i = nn - 2; i >= 0; i--) {
    b[2 * i] = b[2 * i] - a3[i] * b[2 * (i + 1)];
    b[2 * i + 1] = b[2 * i + 1] - a3[i] * b[2 * (i + 1) + 1];
  }
}

/* mkp.f:54-84 */
void o_mkp_coefficients(int nx, int ny, int nz, const double *dx,
                        const double *dy, const double *dz,
                        const double *dy12, double xl, double zl, double *kxz,
                        double *ac) {
  const double pi = 3.14159265358979312;
  const int nh = nx / 2;
  double *kx = malloc(sizeof(double) * (size_t)(nh + 1));
  double *kz = malloc(sizeof(double) * (size_t)(nz + 1));
  for (int i = 0; i <= nh; i++) kx[i] = (double)i / xl * pi * 2.0;
  for (int k = 0; k <= nz; k++) kz[k] = (k <= nz / 2) ? (double)k : (double)(k - nz);
  for (int k = 0; k <= nz - 1; k++) kz[k] = kz[k] / zl * pi * 2.0;
  for (int k = 0; k <= nz; k++)
    for (int i = 0; i <= nh; i++)
      kxz[i + (size_t)(nh + 1) * k] =
          2.0 * (1.0 - cos(kx[i] * dx[i])) / dx[i] / dx[i] +
          2.0 * (1.0 - cos(kz[k] * dz[k])) / dz[k] / dz[k];
  for (int j = 1; j <= ny; j++) {
    ac[j - 1] = 1.0 / dy12[j - 1] / dy[j];
    ac[j - 1 + 2 * (size_t)ny] = 1.0 / dy12[j] / dy[j];
    ac[j - 1 + (size_t)ny] = -ac[j - 1] - ac[j - 1 + 2 * (size_t)ny];
  }
  free(kx);
  free(kz);
}

/* mkp.f:223-260 */
void o_spn(int nx, int ny, int nz, const double *ac, const double *kxz,
           double *phit) {
  const int nh = nx / 2;
  const size_t ty = (size_t)nh + 1, tz = ((size_t)nh + 1) * ((size_t)ny + 2);
#pragma omp parallel
  {
    double *aa = malloc(sizeof(double) * 3 * (size_t)ny);
    double *b = malloc(sizeof(double) * 2 * (size_t)ny);
#pragma omp for schedule(static) collapse(2)
    for (int k = 0; k <= nz - 1; k++)
      for (int i = 0; i <= nh; i++) {
        if ((i + 1) * (k + 1) > 1) {
          for (int j = 1; j <= ny; j++) {
            aa[j - 1] = ac[j - 1];
            aa[j - 1 + ny] = ac[j - 1 + ny] - kxz[i + (size_t)(nh + 1) * k];
            aa[j - 1 + 2 * ny] = ac[j - 1 + 2 * ny];
            size_t o = 2 * ((size_t)i + ty * j + tz * k);
            b[2 * (j - 1)] = phit[o];
            b[2 * (j - 1) + 1] = phit[o + 1];
          }
          aa[ny] = aa[ny] + aa[0];
          aa[0] = 0.0;
          aa[ny - 1 + ny] = aa[ny - 1 + ny] + aa[ny - 1 + 2 * ny];
          aa[ny - 1 + 2 * ny] = 0.0;
          o_ctridiag(aa, b, ny);
          for (int j = 1; j <= ny; j++) {
            size_t o = 2 * ((size_t)i + ty * j + tz * k);
            phit[o] = b[2 * (j - 1)];
            phit[o + 1] = b[2 * (j - 1) + 1];
          }
        }
      }
    free(aa);
    free(b);
  }
}

/* mkp.f:263-290: zero mode by forward march with phi(0) = phi(1) = 0 */
void o_sp0b(int nx, int ny, int nz, const double *ac, double *phit) {
  (void)nz;
  const int nh = nx / 2;
  const size_t ty = (size_t)nh + 1;
  double *d = malloc(sizeof(double) * 2 * (size_t)(ny + 1));
  for (int j = 1; j <= ny; j++) {
    d[2 * j] = phit[2 * (ty * j)];
    d[2 * j + 1] = phit[2 * (ty * j) + 1];
  }
  phit[0] = 0.0; phit[1] = 0.0;
  phit[2 * ty] = 0.0; phit[2 * ty + 1] = 0.0;
  for (int j = 1; j <= ny - 1; j++) {
    double a1 = ac[j - 1], a2 = ac[j - 1 + ny], a3 = ac[j - 1 + 2 * (size_t)ny];
    for (int c = 0; c < 2; c++) {
      double pm = phit[2 * (ty * (j - 1)) + c], p0 = phit[2 * (ty * j) + c];
      phit[2 * (ty * (j + 1)) + c] = (d[2 * j + c] - a1 * pm - a2 * p0) / a3;
    }
  }
  phit[2 * (ty * (ny + 1))] = phit[2 * (ty * ny)];
  phit[2 * (ty * (ny + 1)) + 1] = phit[2 * (ty * ny) + 1];
  free(d);
}

/* mkp.f:293-320 */
void o_phibc(int nx, int ny, int nz, double *phi) {
  GEOM;
  for (int k = 1; k <= nz; k++)
    for (int i = 0; i <= nx; i++) {
      FIELD(phi, i, 0, k) = FIELD(phi, i, 1, k);
      FIELD(phi, i, ny + 1, k) = FIELD(phi, i, ny, k);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      FIELD(phi, i, j, nz + 1) = FIELD(phi, i, j, 1);
      FIELD(phi, i, j, 0) = FIELD(phi, i, j, nz);
    }
  for (int k = 0; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      FIELD(phi, nx + 1, j, k) = FIELD(phi, 1, j, k);
      FIELD(phi, 0, j, k) = FIELD(phi, nx, j, k);
    }
  for (int j = 1; j <= ny; j++) {
    FIELD(phi, 0, j, 0) = FIELD(phi, nx, j, nz);
    FIELD(phi, 0, j, nz + 1) = FIELD(phi, nx, j, 1);
    FIELD(phi, nx + 1, j, 0) = FIELD(phi, 1, j, nz);
    FIELD(phi, nx + 1, j, nz + 1) = FIELD(phi, 1, j, 1);
  }
}

/* mkp.f:9-220 */
void o_mkp(double *p, const double *u_temp, const double *v_temp,
           const double *w_temp, double dt, const double *dx,
           const double *dy, const double *dz, const double *dx12,
           const double *dy12, const double *dz12, double re, int nx, int ny,
           int nz, double *u_ta, double *v_ta, double *w_ta, double *u_ep2,
           double *v_ep2, double *w_ep2, double *divu, double *u, double *v,
           double *w, double xl, double zl) {
  GEOM;
  (void)re;
  const size_t nf = ((size_t)nx + 2) * ((size_t)ny + 2) * ((size_t)nz + 2);
  const int nh = nx / 2;
  double *phi = calloc(nf, sizeof(double));
  double *gx = malloc(sizeof(double) * nf);
  double *gy = malloc(sizeof(double) * nf);
  double *gz = malloc(sizeof(double) * nf);
  double *phit = calloc(2 * ((size_t)nh + 1) * ((size_t)ny + 2) * ((size_t)nz + 2),
                        sizeof(double));
  double *kxz = malloc(sizeof(double) * ((size_t)nh + 1) * ((size_t)nz + 1));
  double *ac = malloc(sizeof(double) * 3 * (size_t)ny);
  o_mkp_coefficients(nx, ny, nz, dx, dy, dz, dy12, xl, zl, kxz, ac);
#pragma omp parallel for schedule(static) collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(phi, i, j, k) =
            ((FIELD(u_temp, i, j, k) - FIELD(u_temp, i - 1, j, k)) / dx[i] +
             (FIELD(v_temp, i, j, k) - FIELD(v_temp, i, j - 1, k)) / dy[j] +
             (FIELD(w_temp, i, j, k) - FIELD(w_temp, i, j, k - 1)) / dz[k]) /
            dt;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) FIELD(phi, i, j, 0) = FIELD(phi, i, j, nz);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) FIELD(phi, 0, j, k) = FIELD(phi, nx, j, k);
  for (int j = 1; j <= ny; j++) FIELD(phi, 0, j, 0) = FIELD(phi, nx, j, nz);
  o_p2s(phi, phit, nx, ny, nz);
  o_spn(nx, ny, nz, ac, kxz, phit);
  o_sp0b(nx, ny, nz, ac, phit);
  o_s2p(phit, phi, nx, ny, nz);
  o_phibc(nx, ny, nz, phi);
#pragma omp parallel for schedule(static) collapse(2)
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        FIELD(gx, i, j, k) = (FIELD(phi, i + 1, j, k) - FIELD(phi, i, j, k)) / dx12[i];
        FIELD(gy, i, j, k) = (FIELD(phi, i, j + 1, k) - FIELD(phi, i, j, k)) / dy12[j];
        FIELD(gz, i, j, k) = (FIELD(phi, i, j, k + 1) - FIELD(phi, i, j, k)) / dz12[k];
      }
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        FIELD(u, i, j, k) = FIELD(u_temp, i, j, k) - dt * FIELD(gx, i, j, k);
        FIELD(v, i, j, k) = FIELD(v_temp, i, j, k) - dt * FIELD(gy, i, j, k);
        FIELD(w, i, j, k) = FIELD(w_temp, i, j, k) - dt * FIELD(gz, i, j, k);
      }
#pragma omp parallel for schedule(static) collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) FIELD(p, i, j, k) = FIELD(phi, i, j, k);
#pragma omp parallel for schedule(static) collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        FIELD(u_ep2, i, j, k) =
            -dt * (FIELD(p, i + 1, j, k) - FIELD(p, i, j, k)) / dx12[i] + FIELD(u_ta, i, j, k);
        FIELD(v_ep2, i, j, k) =
            -dt * (FIELD(p, i, j + 1, k) - FIELD(p, i, j, k)) / dy12[j] + FIELD(v_ta, i, j, k);
        FIELD(w_ep2, i, j, k) =
            -dt * (FIELD(p, i, j, k + 1) - FIELD(p, i, j, k)) / dz12[k] + FIELD(w_ta, i, j, k);
      }
#pragma omp parallel for schedule(static) collapse(2)
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        FIELD(u_ta, i, j, k) = dt * FIELD(gx, i, j, k);
        FIELD(v_ta, i, j, k) = dt * FIELD(gy, i, j, k);
        FIELD(w_ta, i, j, k) = dt * FIELD(gz, i, j, k);
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(divu, i, j, k) = (FIELD(u, i, j, k) - FIELD(u, i - 1, j, k)) / dx[i] +
                               (FIELD(v, i, j, k) - FIELD(v, i, j - 1, k)) / dy[j] +
                               (FIELD(w, i, j, k) - FIELD(w, i, j, k - 1)) / dz[k];
  free(phi); free(gx); free(gy); free(gz); free(phit); free(kxz); free(ac);
}

/* ---------------------------------------------------------- level set */

/* set_levelset.f:1-199.  epsilon = 1.0d0-1 = 0 (set_levelset.f:6), so the
   smoothed sign is phi0/|phi0|; the Heaviside uses pi = 3.1415926535.
   `iterations` is the loop count it = 0..nnn (2001 in the reference). */
void o_set_levelset(int nx, int ny, int nz, const double *dx,
                    const double *dx12, const double *dy12, const double *dz12,
                    const double *phi0, double *phi, const double *x,
                    const double *y, const double *z, double *cellvalue,
                    int iterations) {
  GEOM;
  (void)x; (void)y; (void)z;
  const double dtl = 1.0e-3, epsilon = 1.0 - 1.0, pi = 3.1415926535;
  const size_t nf = ((size_t)nx + 2) * ((size_t)ny + 2) * ((size_t)nz + 2);
  double *tmp = calloc(nf, sizeof(double));
  for (size_t q = 0; q < nf; q++) phi[q] = phi0[q];
  for (int it = 0; it < iterations; it++) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          double c = FIELD(phi, i, j, k);
          double a = (c - FIELD(phi, i - 1, j, k)) / dx12[i - 1];
          double b = (FIELD(phi, i + 1, j, k) - c) / dx[i];
          double cc = (c - FIELD(phi, i, j - 1, k)) / dy12[j - 1];
          double d = (FIELD(phi, i, j + 1, k) - c) / dy12[j];
          double e = (c - FIELD(phi, i, j, k - 1)) / dz12[k - 1];
          double f = (FIELD(phi, i, j, k + 1) - c) / dz12[k];
          double ap = a >= 0.0 ? a : 0.0, am = a >= 0.0 ? 0.0 : a;
          double bp = b >= 0.0 ? b : 0.0, bm = b >= 0.0 ? 0.0 : b;
          double cp = cc >= 0.0 ? cc : 0.0, cm = cc >= 0.0 ? 0.0 : cc;
          double dp = d >= 0.0 ? d : 0.0, dm = d >= 0.0 ? 0.0 : d;
          double ep = e >= 0.0 ? e : 0.0, em = e >= 0.0 ? 0.0 : e;
          double fp = f >= 0.0 ? f : 0.0, fm = f >= 0.0 ? 0.0 : f;
          double p0 = FIELD(phi0, i, j, k);
          double sss = p0 / sqrt(p0 * p0 + epsilon * epsilon);
          double ggg;
          if (p0 > 0.0)
            ggg = sqrt(fmax(ap * ap, bm * bm) + fmax(cp * cp, dm * dm) +
                       fmax(ep * ep, fm * fm)) - 1.0;
          else if (p0 < 0.0)
            ggg = sqrt(fmax(am * am, bp * bp) + fmax(cm * cm, dp * dp) +
                       fmax(em * em, fp * fp)) - 1.0;
          else
            ggg = 0.0;
          FIELD(tmp, i, j, k) = c - dtl * sss * ggg;
        }
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) FIELD(phi, i, j, k) = FIELD(tmp, i, j, k);
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) {
        FIELD(phi, i, j, 0) = FIELD(phi, i, j, 1);
        FIELD(phi, i, j, nz + 1) = FIELD(phi, i, j, nz);
      }
    for (int k = 0; k <= nz + 1; k++)
      for (int i = 0; i <= nx + 1; i++) {
        FIELD(phi, i, 0, k) = FIELD(phi, i, 1, k);
        FIELD(phi, i, ny + 1, k) = FIELD(phi, i, ny, k);
      }
    for (int k = 0; k <= nz + 1; k++)
      for (int j = 0; j <= ny + 1; j++) {
        FIELD(phi, 0, j, k) = FIELD(phi, 1, j, k);
        FIELD(phi, nx + 1, j, k) = FIELD(phi, nx, j, k);
      }
  }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double ph = FIELD(phi, i, j, k);
        if (ph >= dx[i])
          CVAL(i, j, k, 1) = 1.0;
        else if (ph <= -1.0 * dx[i])
          CVAL(i, j, k, 1) = 0.0;
        else
          CVAL(i, j, k, 1) = (sin(pi * ph / 2.0 / dx[i]) + 1.0) / 2.0;
      }
  free(tmp);
}

/* ----------------------------------------------------------- adjoint */

/* adjoint.f:1-126, including its index and factor typos (:45, :72, :79,
   :101, :111), which are part of the operator as written */
void o_set_adj_convect(const double *u, double *u_v, const double *v,
                       double *v_v, const double *w, double *w_v, double re,
                       double dt, int nx, int ny, int nz, const double *dx,
                       const double *dy, const double *dz, const double *dx12,
                       const double *dy12, const double *dz12,
                       const double *uf, const double *vf, const double *wf) {
  GEOM;
  (void)re;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double u1 = -((FIELD(uf, i, j, k) + FIELD(uf, i + 1, j, k)) *
                          (FIELD(u, i, j, k) + FIELD(u, i + 1, j, k)) -
                      (FIELD(uf, i, j, k) + FIELD(uf, i - 1, j, k)) *
                          (FIELD(u, i, j, k) + FIELD(u, i - 1, j, k))) /
                        dx12[i] -
                    FIELD(uf, i, j, k) * (FIELD(u, i - 1, j, k) - FIELD(u, i + 1, j, k)) /
                        dx12[i] * 2.0;
        double u2 = -((FIELD(vf, i + 1, j, k) * dx[i + 1] + FIELD(vf, i, j, k) * dx[i]) *
                          (FIELD(u, i, j + 1, k) + FIELD(u, i, j, k)) -
                      (FIELD(vf, i + 1, j - 1, k) * dx[i + 1] +
                       FIELD(vf, i, j - 1, k) * dx[i]) *
                          (FIELD(u, i, j, k) + FIELD(u, i, j - 1, k))) /
                        dx12[i] / dy[j] -
                    ((FIELD(vf, i + 1, j, k) + FIELD(vf, i + 1, j - 1, k)) * dx[i + 1] +
                     (FIELD(vf, i, j, k) + FIELD(vf, i, j - 1, k)) * dx[i]) /
                        dx12[i] *
                        (-FIELD(v, i + 1, j, k) - FIELD(v, i + 1, j - 1, k) +
                         FIELD(v, i, j, k) + FIELD(v, i, j - 1, k)) /
                        dx12[i] / 2.0;
        double u3 = -((FIELD(wf, i + 1, j, k) * dx[i + 1] + FIELD(wf, i, j, k) * dx[i]) *
                          (FIELD(u, i, j, k + 1) + FIELD(u, i, j, k)) -
                      (FIELD(wf, i + 1, j, k - 1) * dx[i + 1] +
                       FIELD(wf, i, j, k - 1) * dx[i]) *
                          (FIELD(u, i, j, k) + FIELD(u, i, j, k - 1))) /
                        dx12[i] / dz[k] -
                    ((FIELD(wf, i + 1, j, k) + FIELD(wf, i + 1, j, k - 1)) * dx[i + 1] +
                     (FIELD(wf, i, j, k) + FIELD(wf, i, j, k - 1)) * dx[i]) /
                        dx12[i] *
                        (-FIELD(w, i + 1, j, k) - FIELD(w, i + 1, j, k - 1) +
                         FIELD(w, i, j, k) + FIELD(v, i, j, k - 1)) /
                        dx12[i] / 2.0;
        FIELD(u_v, i, j, k) = 0.25 * (-u1 - u2 - u3) * dt + 1.0 * dt;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double v1 = ((FIELD(uf, i, j + 1, k) * dy[j + 1] + FIELD(uf, i, j, k) * dy[j]) *
                         (FIELD(v, i + 1, j, k) + FIELD(v, i, j, k)) -
                     (FIELD(uf, i - 1, j + 1, k) * dy[j + 1] +
                      FIELD(uf, i - 1, j, k) * dy[j]) *
                         (FIELD(v, i, j, k) + FIELD(v, i - 1, j, k))) /
                        dy12[j] / dx[i] -
                    ((FIELD(uf, i, j + 1, k) + FIELD(uf, i - 1, j + 1, k)) * dy[j + 1] +
                     (FIELD(uf, i, j, k) + FIELD(uf, i - 1, j, k)) * dy[j]) /
                        dy12[j] *
                        (-FIELD(u, i, j + 1, k) - FIELD(u, i - 1, j + 1, k) +
                         FIELD(u, i, j, k) + FIELD(u, i - 1, j, k)) /
                        dy12[j] / 2.0;
        double v2 = ((FIELD(vf, i, j, k) + FIELD(vf, i, j + 1, k)) *
                         (FIELD(v, i, j, k) + FIELD(v, i, j + 1, k)) -
                     (FIELD(vf, i, j, k) + FIELD(vf, i, j - 1, k)) *
                         (FIELD(v, i, j, k) + FIELD(v, i, j - 1, k))) /
                        dy12[j] -
                    FIELD(vf, i, j, k) * (FIELD(v, i - 1, j, k) - FIELD(v, i + 1, j, k)) *
                        2.0 / dy12[j];
        double v3 = ((FIELD(wf, i, j + 1, k) * dy[j + 1] + FIELD(wf, i, j, k) * dy[j]) *
                         (FIELD(v, i, j, k + 1) + FIELD(v, i, j, k)) -
                     (FIELD(wf, i, j + 1, k - 1) * dy[j + 1] +
                      FIELD(wf, i, j, k - 1) * dy[j]) *
                         (FIELD(v, i, j, k) + FIELD(v, i, j, k - 1))) /
                        dy12[j] / dz[k] -
                    ((FIELD(wf, i, j, k + 1) + FIELD(wf, i, j + 1, k - 1)) * dy[j + 1] +
                     (FIELD(wf, i, j, k) + FIELD(wf, i, j, k - 1)) * dy[j]) /
                        dy12[j] *
                        (-FIELD(w, i, j + 1, k) - FIELD(w, i, j + 1, k - 1) +
                         FIELD(w, i, j, k) + FIELD(w, i, j, k - 1)) /
                        dy12[j] / 2.0;
        FIELD(v_v, i, j, k) = 0.25 * (-v1 - v2 - v3) * dt;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double w1 = ((FIELD(uf, i, j, k + 1) * dz[k + 1] + FIELD(uf, i, j, k) * dz[k]) *
                         (FIELD(w, i + 1, j, k) + FIELD(w, i, j, k)) -
                     (FIELD(uf, i - 1, j, k + 1) * dz[k + 1] +
                      FIELD(uf, i - 1, j, k) * dz[k]) *
                         (FIELD(w, i, j, k) + FIELD(w, i - 1, j, k))) /
                        dz12[k] / dx[i] -
                    ((FIELD(uf, i, j, k + 1) + FIELD(uf, i - 1, j, k + 1)) * dz[k + 1] +
                     (FIELD(uf, i, j, k) + FIELD(uf, i - 1, j, k)) * dz[k]) /
                        dz12[k] *
                        (-FIELD(u, i, j, k + 1) - FIELD(u, i - 1, j, k + 1) +
                         FIELD(u, i, j, k) + FIELD(u, i - 1, j, k)) /
                        dz12[k] * 2.0;
        double w2 = ((FIELD(vf, i, j, k + 1) * dz[k + 1] + FIELD(vf, i, j, k) * dz[k]) *
                         (FIELD(w, i, j + 1, k) + FIELD(w, i, j, k)) -
                     (FIELD(vf, i, j - 1, k + 1) * dz[k + 1] +
                      FIELD(vf, i, j - 1, k) * dz[k]) *
                         (FIELD(w, i, j, k) + FIELD(w, i, j - 1, k))) /
                        dz12[k] / dy[j] -
                    ((FIELD(vf, i, j, k + 1) + FIELD(vf, i, j - 1, k + 1)) * dz[k + 1] +
                     (FIELD(vf, i, j, k) + FIELD(vf, i, j - 1, k)) * dz[k]) /
                        dz12[k] *
                        (-FIELD(v, i, j, k + 1) - FIELD(v, i, j - 1, k + 1) +
                         FIELD(v, i, j, k) + FIELD(v, i, j - 1, k)) /
                        dz12[k] * 2.0;
        double w3 = ((FIELD(wf, i, j, k) + FIELD(wf, i, j, k + 1)) *
                         (FIELD(w, i, j, k) + FIELD(w, i, j, k + 1)) -
                     (FIELD(wf, i, j, k) + FIELD(wf, i, j, k - 1)) *
                         (FIELD(w, i, j, k) + FIELD(w, i, j, k - 1))) /
                        dz12[k] -
                    FIELD(wf, i, j, k) * (FIELD(w, i, j, k - 1) - FIELD(w, i, j, k + 1)) *
                        2.0 / dz12[k];
        FIELD(w_v, i, j, k) = 0.25 * (-w1 - w2 - w3) * dt;
      }
}

/* adjoint.f:128-181 */
void o_set_adj_convection_for_scalar(
    const double *u, const double *v, const double *w, double re, double dt,
    int nx, int ny, int nz, const double *dx, const double *dy,
    const double *dz, const double *dx12, const double *dy12,
    const double *dz12, const double *t, double *u_vt, const double *rhocp,
    const double *uf, const double *vf, const double *wf) {
  GEOM;
  (void)u; (void)v; (void)w; (void)re;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double rc = FIELD(rhocp, i, j, k), tc = FIELD(t, i, j, k);
        double ut = (-((FIELD(rhocp, i - 1, j, k) * FIELD(t, i - 1, j, k) * dx[i] +
                        rc * tc * dx[i - 1]) *
                       FIELD(uf, i - 1, j, k)) /
                         2.0 / dx12[i - 1] +
                     ((rc * tc * dx[i + 1] +
                       FIELD(rhocp, i + 1, j, k) * FIELD(t, i + 1, j, k) * dx[i]) *
                      FIELD(uf, i, j, k)) /
                         2.0 / dx12[i]) /
                    dx[i];
        double vt = (-((FIELD(rhocp, i, j - 1, k) * FIELD(t, i, j - 1, k) * dy[j] +
                        rc * tc * dy[j - 1]) *
                       FIELD(vf, i, j - 1, k)) /
                         2.0 / dy12[j - 1] +
                     ((rc * tc * dy[j + 1] +
                       FIELD(rhocp, i, j + 1, k) * FIELD(t, i, j + 1, k) * dy[j]) *
                      FIELD(vf, i, j, k)) /
                         2.0 / dy12[j]) /
                    dy[j];
        double wt = (-((FIELD(rhocp, i, j, k - 1) * FIELD(t, i, j, k - 1) * dz[k] +
                        rc * tc * dz[k - 1]) *
                       FIELD(wf, i, j, k - 1)) /
                         2.0 / dz12[k - 1] +
                     ((rc * tc * dz[k + 1] +
                       FIELD(rhocp, i, j, k + 1) * FIELD(t, i, j, k + 1) * dz[k]) *
                      FIELD(wf, i, j, k)) /
                         2.0 / dz12[k]) /
                    dz[k];
        FIELD(u_vt, i, j, k) = (ut + vt + wt) * dt / rc + 1.0 * dt;
      }
}

/* adjoint.f:185-211 */
void o_set_adj_ns_scalar(double dt, int nx, int ny, int nz,
                         const double *dx12, const double *dy12,
                         const double *dz12, const double *t, double *u_s,
                         double *v_s, double *w_s, const double *uf,
                         const double *vf, const double *wf,
                         const double *tf) {
  GEOM;
  (void)uf; (void)vf; (void)wf;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double f = FIELD(tf, i, j, k);
        FIELD(u_s, i, j, k) =
            dt * f * (FIELD(t, i + 1, j, k) - FIELD(t, i - 1, j, k)) / dx12[i] / 2.0;
        FIELD(v_s, i, j, k) =
            dt * f * (FIELD(t, i, j + 1, k) - FIELD(t, i, j - 1, k)) / dy12[j] / 2.0;
        FIELD(w_s, i, j, k) =
            dt * f * (FIELD(t, i, j, k + 1) - FIELD(t, i, j, k - 1)) / dz12[k] / 2.0;
      }
}

/* ----------------------------------------------------------- one step */

/* nse_dominik.f:290-448 */
void o_time_step(o_state *s, int l) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  GEOM;
  double cons_cell;
  o_set_bc(s->u, s->v, s->w, s->p, nx, ny, nz, s->t);
  o_set_convect(s->u, s->u_v, s->v, s->v_v, s->w, s->w_v, s->dt, nx, ny, nz,
                s->dx, s->dy, s->dz, s->dx12, s->dy12, s->dz12);
  o_set_convection_for_scalar(s->u, s->v, s->w, s->dt, nx, ny, nz, s->dx,
                              s->dy, s->dz, s->dx12, s->dy12, s->dz12, s->t,
                              s->u_vt, s->rhocp);
  o_set_conduction(s->u, s->u_d, s->v, s->v_d, s->w, s->w_d, s->re, s->dt, nx,
                   ny, nz, s->dx, s->dy, s->dz, s->dx12, s->dy12, s->dz12);
  o_set_diffusion_for_scalar(s->re, s->dt, nx, ny, nz, s->dx, s->dy, s->dz,
                             s->dx12, s->dy12, s->dz12, s->t, s->u_dt, s->sc,
                             s->thc, s->rhocp);
  o_set_immersed_boundary(s->eta, nx, ny, nz, s->dt, s->u_l, s->v_l, s->w_l,
                          s->dx, s->dy, s->dz, s->t_l, s->cellvalue, s->u_ta,
                          s->v_ta, s->w_ta, l, &cons_cell);
  o_set_explicit(s->u_v, s->u_pre, s->u_d, s->u_a, s->v_v, s->v_pre, s->v_d,
                 s->v_a, s->w_v, s->w_pre, s->w_d, s->w_a, s->u_l, s->v_l,
                 s->w_l, nx, ny, nz, l);
  o_set_explicit_for_tempreture(s->u_vt, s->u_dt, s->u_t, s->t_pre, s->t_l,
                                nx, ny, nz, l);
  o_set_bc(s->u_a, s->v_a, s->w_a, s->p, nx, ny, nz, s->t);
  o_t_divide(s->u_t, s->t, s->cellvalue, s->eta, s->dt, nx, ny, nz);
  o_solve_u_temp(s->u_a, s->v_a, s->w_a, nx, ny, nz, s->dx, s->dy, s->dz,
                 s->dx12, s->dy12, s->dz12, s->re, s->dt, s->eta,
                 s->cellvalue, s->u, s->v, s->w);
  const size_t nf = sz * ((size_t)nz + 2);
  memcpy(s->u_temp, s->u_a, nf * sizeof(double));
  memcpy(s->v_temp, s->v_a, nf * sizeof(double));
  memcpy(s->w_temp, s->w_a, nf * sizeof(double));
  o_solve_t(s->u_t, nx, ny, nz, s->dx, s->dy, s->dz, s->dx12, s->dy12,
            s->dz12, s->re, s->sc, s->dt, s->rhocp, s->thc, s->cellvalue,
            s->eta);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) FIELD(s->t, i, j, k) = FIELD(s->u_t, i, j, k);
  o_set_bc(s->u_temp, s->v_temp, s->w_temp, s->p, nx, ny, nz, s->t);
  o_mkp(s->p, s->u_temp, s->v_temp, s->w_temp, s->dt, s->dx, s->dy, s->dz,
        s->dx12, s->dy12, s->dz12, s->re, nx, ny, nz, s->u_ta, s->v_ta,
        s->w_ta, s->u_ep2, s->v_ep2, s->w_ep2, s->divu, s->u, s->v, s->w,
        s->xl, s->zl);
}

/* ------------------------------------------------------- diagnostics */

/* nse_dominik.f:2519-2601.  Each ratio is (x dt) / d as the reference
   forms it; the three maxima start at 0 and take a value only when it is
   larger (a negative or NaN ratio never enters), and the result is their
   maximum. */
void o_check_courant(double dt, const double *dx, const double *dy, const double *dz, int nx,
                     int ny, int nz, const double *u, const double *v, const double *w,
                     double *courant) {
  GEOM;
  double mx = 0.0, my = 0.0, mz = 0.0;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double cx = FIELD(u, i, j, k) * dt / dx[i];
        const double cy = FIELD(v, i, j, k) * dt / dy[j];
        const double cz = FIELD(w, i, j, k) * dt / dz[k];
        if (mx < cx) mx = cx;
        if (my < cy) my = cy;
        if (mz < cz) mz = cz;
      }
  /* max(courant_max_x, courant_max_y, courant_max_z) (:2591) */
  double c = mx;
  if (my > c) c = my;
  if (mz > c) c = mz;
  *courant = c;
}

/* nse_dominik.f:2797-2872.  The three velocity integrals accumulate in one
   k, j, i loop (:2814-2822); vel_dif leaves out w (:2829).  T_all is summed
   over j, i after the k loop has ended, so k = nz + 1 there, the top halo
   plane (:2838-2843).  The largest |T - T_before| keeps its first position
   in k, j, i order (strict >, :2854-2864). */
void o_monitor(int nx, int ny, int nz, int l, const double *t, const double *dx,
               const double *dy, const double *dz, double *dif, double *t_all, double *t_dif,
               const double *t_before, const double *u, const double *v, const double *w,
               double *vel_dif, double *u_dif, double *v_dif, double *w_dif,
               double *u_int_before, double *v_int_before, double *w_int_before, double *t_b,
               double dt, int *loc) {
  GEOM;
  (void)dt;
  if (l == 1) {
    *t_b = 0.0;
    *u_int_before = 0.0;
    *v_int_before = 0.0;
    *w_int_before = 0.0;
  }
  double u_int = 0.0, v_int = 0.0, w_int = 0.0;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        u_int = u_int + fabs(FIELD(u, i, j, k)) * dx[i];
        v_int = v_int + fabs(FIELD(v, i, j, k)) * dy[j];
        w_int = w_int + fabs(FIELD(w, i, j, k)) * dz[k];
      }
  *u_dif = fabs((u_int - *u_int_before) / u_int);
  *v_dif = fabs((v_int - *v_int_before) / v_int);
  *w_dif = fabs((w_int - *w_int_before) / w_int);
  *vel_dif = (*u_dif >= *v_dif) ? *u_dif : *v_dif;
  *u_int_before = u_int;
  *v_int_before = v_int;
  *w_int_before = w_int;
  const int kt = nz + 1;
  double ta = 0.0;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) ta = ta + FIELD(t, i, j, kt) * dx[i] * dy[j] * dz[kt];
  *t_all = ta;
  *dif = fabs((ta - *t_b) / ta);
  *t_b = ta;
  double td = 0.0;
  loc[0] = loc[1] = loc[2] = 0;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double e = fabs(FIELD(t, i, j, k) - FIELD(t_before, i, j, k));
        if (e > td) {
          td = e;
          loc[0] = i;
          loc[1] = j;
          loc[2] = k;
        }
      }
  *t_dif = td;
}

/* nse_dominik.f:2875-2892: sum over k, j of |u(nx,j,k) - (-Re/2 y (y - yl))|,
   the analytic value formed as -(((Re/2) y(j)) (y(j) - yl)). */
void o_check_convergence(const double *y, double yl, const double *u, double re, int nx,
                         int ny, int nz, double *error_sum) {
  GEOM;
  double e = 0.0;
  const int i = nx;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      const double a = -(re / 2.0 * y[j] * (y[j] - yl));
      e = e + fabs(FIELD(u, i, j, k) - a);
    }
  *error_sum = e;
}

/* The analytic shape union (no reference routine; see oracle.h). */
void o_levelset_shapes(int nx, int ny, int nz, const double *x, const double *y,
                       const double *z, double xl, double zl, int nshapes,
                       const double *shapes, double *phi0) {
  const long sy = nx + 2, sz = (long)(nx + 2) * (ny + 2);
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) {
        const long o = i + sy * j + sz * k;
        if (i == 0 || j == 0 || k == 0 || i == nx + 1 || j == ny + 1 || k == nz + 1) {
          phi0[o] = 0.0;
          continue;
        }
        double best = -1.0;
        for (int q = 0; q < nshapes; q++) {
          const double *p = shapes + 10 * (long)q;
          double w = p[6], qx = p[7], qy = p[8], qz = p[9];
          const double nq = sqrt(w * w + qx * qx + qy * qy + qz * qz);
          w /= nq;
          qx /= nq;
          qy /= nq;
          qz /= nq;
          const double r0 = 1.0 - 2.0 * (qy * qy + qz * qz), r1 = 2.0 * (qx * qy - qz * w),
                       r2 = 2.0 * (qx * qz + qy * w), r3 = 2.0 * (qx * qy + qz * w),
                       r4 = 1.0 - 2.0 * (qx * qx + qz * qz), r5 = 2.0 * (qy * qz - qx * w),
                       r6 = 2.0 * (qx * qz - qy * w), r7 = 2.0 * (qy * qz + qx * w),
                       r8 = 1.0 - 2.0 * (qx * qx + qy * qy);
          double dx = x[i] - p[0], dy = y[j] - p[1], dz = z[k] - p[2];
          if (dx > 0.5 * xl) dx -= xl;
          if (dx < -0.5 * xl) dx += xl;
          if (dz > 0.5 * zl) dz -= zl;
          if (dz < -0.5 * zl) dz += zl;
          const double b0 = (r0 * dx + r3 * dy + r6 * dz) * (1.0 / p[3]);
          const double b1 = (r1 * dx + r4 * dy + r7 * dz) * (1.0 / p[4]);
          const double b2 = (r2 * dx + r5 * dy + r8 * dz) * (1.0 / p[5]);
          const double amin = fmin(p[3], fmin(p[4], p[5]));
          const double v = amin * (1.0 - sqrt(b0 * b0 + b1 * b1 + b2 * b2));
          best = q == 0 ? v : fmax(best, v);
        }
        phi0[o] = best;
      }
}

/* ------------------------------------------- iterative Poisson (dead) */

/* implicit_px/py/pz (nse_dominik.f:2270-2457): rows 0..n+1 of the dense
   matrix the reference builds, kept as its three bands -- ghost rows
   p0 - p1 = ft, p(n+1) - p(n) = fb, interior rows (blo, -bsum, bhi) --
   eliminated as the reference eliminates them; f(1..n) in/out. */
static void p_line(double *f, int n, double blo, double bhi, double bsum,
                   double ft, double fb) {
  int m = n + 1;
  double *lo = calloc((size_t)4 * (m + 1), sizeof(double));
  double *di = lo + (m + 1), *up = di + (m + 1), *u = up + (m + 1);
  di[0] = 1.0;
  up[0] = -1.0;
  di[n + 1] = 1.0;
  lo[n + 1] = -1.0;
  for (int i = 1; i <= n; i++) {
    lo[i] = blo;
    di[i] = -(bsum);
    up[i] = bhi;
  }
  for (int i = 1; i <= n; i++) u[i] = f[i];
  u[0] = ft;
  u[n + 1] = fb;
  ul_eliminate(lo, di, up, u, m);
  for (int i = 1; i <= n; i++) f[i] = u[i];
  free(lo);
}

void o_linebyline(int nx, int ny, int nz, const double *u_temp,
                  const double *v_temp, const double *w_temp, const double *dx,
                  const double *dy, const double *dz, const double *dx12,
                  const double *dy12, const double *dz12, double dt, double *p) {
  GEOM;
  const int nm = (nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz)) + 2;
  double *q = calloc((size_t)(nx + 2) * (ny + 2) * (nz + 2), sizeof(double));
  double *f = calloc((size_t)nm, sizeof(double));
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        FIELD(q, i, j, k) = ((FIELD(u_temp, i, j, k) - FIELD(u_temp, i - 1, j, k)) / dx[i] +
                             (FIELD(v_temp, i, j, k) - FIELD(v_temp, i, j - 1, k)) / dy[j] +
                             (FIELD(w_temp, i, j, k) - FIELD(w_temp, i, j, k - 1)) / dz[k]) /
                            dt;
  double bx1 = 0, bx2 = 0, by1 = 0, by2 = 0, bz1 = 0, bz2 = 0;
#define PB(i, j, k)                        \
  bx1 = 1.0 / dx12[i] / dx[i];             \
  bx2 = 1.0 / dx12[(i) - 1] / dx[i];       \
  by1 = 1.0 / dy12[j] / dy[j];             \
  by2 = 1.0 / dy12[(j) - 1] / dy[j];       \
  bz1 = 1.0 / dz12[k] / dz[k];             \
  bz2 = 1.0 / dz12[(k) - 1] / dz[k]
#define BSUM (bx1 + bx2 + by1 + by2 + bz1 + bz2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      for (int i = 1; i <= nx; i++) {
        PB(i, j, k);
        f[i] = -by1 * FIELD(p, i, j + 1, k) - by2 * FIELD(p, i, j - 1, k) -
               bz1 * FIELD(p, i, j, k + 1) - bz2 * FIELD(p, i, j, k - 1) + FIELD(q, i, j, k);
      }
      p_line(f, nx, bx2, bx1, BSUM, 0.0, 0.0);
      for (int i = 1; i <= nx; i++) FIELD(p, i, j, k) = f[i];
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      for (int j = 1; j <= ny; j++) {
        PB(i, j, k);
        f[j] = -bx1 * FIELD(p, i + 1, j, k) - bx2 * FIELD(p, i - 1, j, k) -
               bz1 * FIELD(p, i, j, k + 1) - bz2 * FIELD(p, i, j, k - 1) + FIELD(q, i, j, k);
      }
      p_line(f, ny, by2, by1, BSUM, 0.0, 0.0);
      for (int j = 1; j <= ny; j++) FIELD(p, i, j, k) = f[j];
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      for (int k = 1; k <= nz; k++) {
        PB(i, j, k);
        f[k] = -bx1 * FIELD(p, i + 1, j, k) - bx2 * FIELD(p, i - 1, j, k) -
               by1 * FIELD(p, i, j + 1, k) - by2 * FIELD(p, i, j - 1, k) + FIELD(q, i, j, k);
      }
      p_line(f, nz, bz2, bz1, BSUM, 0.0, 0.0);
      for (int k = 1; k <= nz; k++) FIELD(p, i, j, k) = f[k];
    }
#undef PB
#undef BSUM
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 1; i++) {
      FIELD(p, i, j, 0) = FIELD(p, i, j, 1);
      FIELD(p, i, j, nz + 1) = FIELD(p, i, j, nz);
    }
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++) {
      FIELD(p, 0, j, k) = FIELD(p, 1, j, k);
      FIELD(p, nx + 1, j, k) = FIELD(p, nx, j, k);
    }
  for (int k = 0; k <= nz + 1; k++)
    for (int i = 0; i <= nx + 1; i++) {
      FIELD(p, i, 0, k) = FIELD(p, i, 1, k);
      FIELD(p, i, ny + 1, k) = FIELD(p, i, ny, k);
    }
  const double p_st = FIELD(p, 1, 1, 1);
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= nx + 1; i++) FIELD(p, i, j, k) = FIELD(p, i, j, k) - p_st;
  free(q);
  free(f);
}

void o_solve_p(double *p, const double *u_temp, const double *v_temp,
               const double *w_temp, double dt, const double *dx,
               const double *dy, const double *dz, const double *dx12,
               const double *dy12, const double *dz12, int nx, int ny, int nz,
               double *res, double *res_max, double *res_int, double *res_var,
               int max_iter, int *iters) {
  GEOM;
  *iters = 0;
  for (int lll = 1; lll <= max_iter; lll++) {
    o_linebyline(nx, ny, nz, u_temp, v_temp, w_temp, dx, dy, dz, dx12, dy12, dz12, dt, p);
    *iters = lll;
    *res = 0.0;
    *res_max = 0.0;
    *res_int = 0.0;
    *res_var = 0.0;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const double div = (FIELD(u_temp, i, j, k) - FIELD(u_temp, i - 1, j, k)) / dx[i] +
                             (FIELD(v_temp, i, j, k) - FIELD(v_temp, i, j - 1, k)) / dy[j] +
                             (FIELD(w_temp, i, j, k) - FIELD(w_temp, i, j, k - 1)) / dz[k];
          const double pphi =
              ((FIELD(p, i + 1, j, k) - FIELD(p, i, j, k)) / dx12[i] -
               (FIELD(p, i, j, k) - FIELD(p, i - 1, j, k)) / dx12[i - 1]) / dx[i] +
              ((FIELD(p, i, j + 1, k) - FIELD(p, i, j, k)) / dy12[j] -
               (FIELD(p, i, j, k) - FIELD(p, i, j - 1, k)) / dy12[j - 1]) / dy[j] +
              ((FIELD(p, i, j, k + 1) - FIELD(p, i, j, k)) / dz12[k] -
               (FIELD(p, i, j, k) - FIELD(p, i, j, k - 1)) / dz12[k - 1]) / dz[k];
          *res = fabs(pphi * dt - div);
          if (*res >= *res_max) *res_max = *res;
          *res_int = *res_int + (pphi * dt - div) * dx[i] * dy[j] * dz[k];
          *res_var = *res_var + (*res * *res) * dx[i] * dy[j] * dz[k];
        }
    if (*res_max < 1.0e-005) break;
  }
}
